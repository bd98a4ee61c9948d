// Option parsing and validation. Keys, defaults, error messages and the
// configure_from rules follow the reference (src/fasteig_options.cpp:19-195).
#include "fasteig_options.hpp"

#include <algorithm>
#include <cctype>
#include <limits>
#include <sstream>

namespace {
std::string lower(std::string_view s) {
  std::string out(s.begin(), s.end());
  std::transform(out.begin(), out.end(), out.begin(),
                 [](unsigned char c) { return static_cast<char>(std::tolower(c)); });
  return out;
}
std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}
}  // namespace

FASTEIG_INT fasteig_options::parse_int(std::string_view key, std::string_view v) {
  try {
    size_t pos = 0;
    const int value = std::stoi(std::string(v), &pos);
    if (pos != v.size()) throw std::invalid_argument("");
    return static_cast<FASTEIG_INT>(value);
  } catch (...) {
    throw std::invalid_argument("Invalid integer for key '" + std::string(key) + "': '" + std::string(v) + "'");
  }
}

FASTEIG_DOUBLE fasteig_options::parse_double(std::string_view key, std::string_view v) {
  try {
    size_t pos = 0;
    const double value = std::stod(std::string(v), &pos);
    if (pos != v.size()) throw std::invalid_argument("");
    return value;
  } catch (...) {
    throw std::invalid_argument("Invalid double for key '" + std::string(key) + "': '" + std::string(v) + "'");
  }
}

bool fasteig_options::parse_bool(std::string_view key, std::string_view v) {
  const std::string s = lower(v);
  if (s == "1" || s == "true" || s == "yes" || s == "on") return true;
  if (s == "0" || s == "false" || s == "no" || s == "off") return false;
  throw std::invalid_argument("Invalid boolean for key '" + std::string(key) + "': '" + std::string(v) + "'");
}

bool fasteig_options::set(std::string_view key, std::string_view value) {
  const std::string k = lower(key);
  if (k == "n") n_ = parse_int(key, value);
  else if (k == "nev") nev_ = parse_int(key, value);
  else if (k == "maxiter") maxiter_ = parse_int(key, value);
  else if (k == "maxdim") maxdim_ = parse_int(key, value);
  else if (k == "maxblks") maxblks_ = parse_int(key, value);
  else if (k == "degree") degree_ = parse_int(key, value);
  else if (k == "which") which_ = std::string(value);
  else if (k == "solver") solver_ = std::string(value);
  else if (k == "shift") shift_ = parse_double(key, value);
  else if (k == "blksize") blksize_ = parse_int(key, value);
  else if (k == "tol") tol_ = parse_double(key, value);
  else if (k == "seed") {
    const long long s = std::stoll(std::string(value));
    if (s < 0) seed_.reset();
    else if (s > static_cast<long long>(std::numeric_limits<unsigned int>::max()))
      throw std::invalid_argument("seed out of range for unsigned int");
    else seed_ = static_cast<unsigned int>(s);
  } else if (k == "verbosity") {
    verbosity_ = static_cast<int>(parse_int(key, value));
    if (verbosity_ <= 1) verbosity_ = 1;
    if (verbosity_ >= 2) verbosity_ = 2;
  } else if (k == "print_every") print_every_ = parse_int(key, value);
  else if (k == "verbose") {
    (void)parse_bool(key, value);
    verbosity_ = 1;
  } else if (k == "logprefix" || k == "logfile" || k == "output_prefix") logprefix_ = std::string(value);
  else if (k == "output_format") output_format_ = std::string(value);
  else if (k == "gpu_backend") gpu_backend_ = parse_bool(key, value);
  else if (k == "precision") {
    const std::string v = lower(value);
    if (v != "fp64" && v != "mixed") throw std::invalid_argument("precision must be 'fp64' or 'mixed'");
    precision_ = v;
  } else if (k == "gram") {
    const std::string v = lower(value);
    if (v != "native" && v != "bf16x3" && v != "bf16x6" && v != "auto")
      throw std::invalid_argument("gram must be 'native', 'bf16x3', 'bf16x6' or 'auto'");
    gram_ = v;
  } else if (k == "format") {
    const std::string v = lower(value);
    if (v != "csr" && v != "sell" && v != "blocked" && v != "tiled" && v != "auto")
      throw std::invalid_argument("format must be 'csr', 'sell', 'blocked', 'tiled' or 'auto'");
    format_ = v;
  } else if (k == "cheb_bounds") {
    const std::string v = lower(value);
    if (v != "amplification" && v != "ritz") throw std::invalid_argument("cheb_bounds must be 'amplification' or 'ritz'");
    cheb_bounds_ = v;
  } else if (k == "device") device_ = static_cast<int>(parse_int(key, value));
  else return false;
  return true;
}

std::string fasteig_options::get(std::string_view key) const {
  const std::string k = lower(key);
  if (k == "n") return std::to_string(n_);
  if (k == "nev") return std::to_string(nev_);
  if (k == "maxiter") return std::to_string(maxiter_);
  if (k == "maxdim") return std::to_string(maxdim_);
  if (k == "maxblks") return std::to_string(maxblks_);
  if (k == "degree") return std::to_string(degree_);
  if (k == "which") return which_;
  if (k == "solver") return solver_;
  if (k == "shift") return std::to_string(shift_);
  if (k == "blksize") return std::to_string(blksize_);
  if (k == "tol") return std::to_string(tol_);
  if (k == "seed") return seed_.has_value() ? std::to_string(*seed_) : "-1";
  if (k == "verbosity") return std::to_string(verbosity_);
  if (k == "print_every") return std::to_string(print_every_);
  if (k == "verbose") return verbosity_ > 0 ? "true" : "false";
  if (k == "logprefix" || k == "logfile" || k == "output_prefix") return logprefix_;
  if (k == "output_format") return output_format_;
  if (k == "gpu_backend") return gpu_backend_ ? "true" : "false";
  if (k == "precision") return precision_;
  if (k == "gram") return gram_;
  if (k == "format") return format_;
  if (k == "cheb_bounds") return cheb_bounds_;
  if (k == "device") return std::to_string(device_);
  throw std::out_of_range("Unknown option key: " + std::string(key));
}

bool fasteig_options::apply(const std::unordered_map<std::string, std::string>& kv, bool strict) {
  for (const auto& [k, v] : kv) {
    const bool ok = set(k, v);
    if (!ok && strict) throw std::invalid_argument("Unknown option key: " + k);
  }
  return true;
}

void fasteig_options::apply_string(const std::string& s) {
  std::stringstream ss(s);
  std::string item;
  while (std::getline(ss, item, ';')) {
    item = trim(item);
    if (item.empty()) continue;
    const size_t eq = item.find('=');
    if (eq == std::string::npos) throw std::invalid_argument("option '" + item + "' is not key=value");
    const std::string k = trim(item.substr(0, eq)), v = trim(item.substr(eq + 1));
    if (!set(k, v)) throw std::invalid_argument("Unknown option key: " + k);
  }
}

void fasteig_options::configure_from(const fasteig_params& p) {
  logprefix_ = p.logprefix;
  output_format_ = p.output_format;
  n_ = p.n;
  if (p.nev > 0) nev_ = p.nev;
  if (p.maxiter > 0) maxiter_ = p.maxiter;
  if (p.maxdim > 0) maxdim_ = p.maxdim;
  if (p.degree > 0) degree_ = p.degree;
  if (p.blksize > 0) blksize_ = p.blksize;
  if (p.maxblks > 0) maxblks_ = p.maxblks;
  which_ = p.which;
  solver_ = p.solver;
  shift_ = p.shift;
  tol_ = p.tol;
  if (p.seed >= 0) seed_ = static_cast<unsigned int>(p.seed);
  if (p.verbosity >= 0) verbosity_ = p.verbosity;
  if (verbosity_ < 1) verbosity_ = 1;
  if (verbosity_ > 2) verbosity_ = 2;
  if (p.print_every > 0) print_every_ = p.print_every;
  if (maxblks_ <= 0) maxblks_ = std::max<FASTEIG_INT>(10, nev_);
  set("precision", p.precision);
  set("gram", p.gram);
  set("format", p.format);
  if (!p.cheb_bounds.empty()) set("cheb_bounds", p.cheb_bounds);
}

void fasteig_options::validate() const {
  if (n_ < 0) throw std::invalid_argument("options.n must be >= 0");
  if (nev_ <= 0) throw std::invalid_argument("options.nev must be > 0");
  if (maxiter_ <= 0) throw std::invalid_argument("options.maxiter must be > 0");
  if (tol_ <= 0) throw std::invalid_argument("options.tol must be > 0");
  if (degree_ <= 0) throw std::invalid_argument("options.degree must be > 0");
  if (maxblks_ < 0) throw std::invalid_argument("options.maxblks must be >= 0");
  if (blksize_ <= 0) throw std::invalid_argument("options.blksize must be > 0");
  if (verbosity_ < 1 || verbosity_ > 2) throw std::invalid_argument("options.verbosity must be 1 or 2");
  if (print_every_ < 0) throw std::invalid_argument("options.print_every must be >= 0");
  if ((gram_ == "bf16x3" || gram_ == "bf16x6") && precision_ != "mixed")
    throw std::invalid_argument("options.gram=" + gram_ + " requires precision=mixed");
}
