// Internal definitions shared by the gfx950 kernel sources of libfe_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/fe_hip.h"

namespace fe {

// thread-local error message behind fe_last_error()
void set_error(const std::string& msg);

// Growable device scratch owned by a context (split-K partials, reductions,
// eigensolver work). Grown with hipMalloc only when a call needs more; never
// shrunk until the context dies.
struct Scratch {
  void* ptr = nullptr;
  size_t bytes = 0;
};

// Stream-ordered reuse of fe_alloc'd blocks: fe_free parks a block here
// instead of synchronizing and calling hipFree; a later fe_alloc of a size
// within 25% takes it back. All work of a context runs on its one stream, so
// a parked block is reused only by work queued after its last use. The cache
// is bounded (an eighth of HBM) and emptied when hipMalloc runs out.
struct BlockCache {
  std::mutex mu;
  std::multimap<size_t, void*> parked;       // size -> block
  std::unordered_map<void*, size_t> live;    // block -> rounded size
  size_t parked_bytes = 0;
};

}  // namespace fe

struct fe_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  int cus = 0;
  size_t hbm = 0;
  char name[128] = {0};
  fe::Scratch ws[7];           // independent scratch slots (4: split coefficients of the BF16 products, 5: small results, 6: column-sum segments)
  fe::BlockCache cache;        // fe_alloc / fe_free reuse
  int* info_dev = nullptr;     // small device status words
  void* gram_static = nullptr; // device address of fe_gram_static's fixed scratch on this device
  hipEvent_t events[16] = {};
  // captured launch sequences (the tridiagonalization of fe_syev, the
  // blocked Cholesky and triangular inverse), keyed by the routine, its sizes
  // and the device addresses the launches were recorded with
  struct SeqGraph {
    int64_t key[8] = {};
    hipGraphExec_t exec = nullptr;
  };
  std::vector<SeqGraph> seq_graphs;
};

namespace fe {
// Runs `launches` (kernel launches on ctx->stream only) as a graph captured on
// the first call with this key and replayed afterwards; when the stream is
// being captured by the caller, or capture is not wanted, launches directly.
// FE_GRAPHS=0 in the environment (read once per process) turns the capture
// off: the same launches then go out one by one, for tools that trace kernel
// launches but cannot follow a captured stream.
int run_captured(fe_ctx ctx, const int64_t (&key)[8], const std::function<void()>& launches);
}  // namespace fe

struct fe_csr_s {
  fe_ctx ctx = nullptr;
  int n = 0;
  int64_t nnz = 0;
  int* rowptr = nullptr;
  int* colind = nullptr;
  double* vals64 = nullptr;
  float* vals32 = nullptr;
  // SELL-C-sigma copy (optional)
  int sell_C = 0, sell_sigma = 0;
  int64_t sell_slices = 0;
  int64_t sell_padded = 0;     // stored entries including padding
  int64_t* sell_ptr = nullptr; // slice start offsets (slices + 1)
  int* sell_len = nullptr;     // slice widths (max row length in the slice)
  int* sell_perm = nullptr;    // sorted position -> original row
  int* sell_col = nullptr;     // column-major within a slice: [slice][w][C]
  double* sell_v64 = nullptr;
  float* sell_v32 = nullptr;
  int use_sell = 0;
  // row-block copy (optional): blocks of bg_rb consecutive rows, each with the
  // sorted union of its rows' columns (bg_ucol, offsets bg_bptr) and bg_rb
  // values per union entry (bg_v64 / bg_v32, zero where a row lacks the column)
  int bg_rb = 0;
  int64_t bg_nblk = 0;
  int64_t bg_nu = 0;
  int64_t* bg_bptr = nullptr;
  int* bg_ucol = nullptr;
  double* bg_v64 = nullptr;
  float* bg_v32 = nullptr;
  double bg_reuse = 0.0;      // nonzeros / union entries (gathers saved)
  int* bg_sched = nullptr;     // optional processing order of the row blocks
  int64_t bg_plane = 0, bg_strip = 0;
  int use_bg = 0;
  // tiled copy (optional): tiles of 16 rows taken in pencils of consecutive
  // tiles; a workgroup keeps the X rows its tile needs in LDS slots and loads
  // only the rows the next tile adds (fe_spmm_tile.hip)
  int64_t tl_ntiles = 0, tl_npencils = 0, tl_loads = 0;
  int tl_slots = 0;            // LDS row slots per workgroup; slot tl_slots is the zero row
  int64_t* tl_pencil = nullptr;  // first tile of each pencil (npencils + 1)
  int* tl_rows = nullptr;        // tl_trows rows per tile (-1: none)
  // entries: every row of a tile holds tl_maxe[tile] of them (its nonzeros in
  // CSR order, then the zero slot with value 0), rows one after another
  int64_t* tl_ebase = nullptr;   // first entry of each tile (ntiles + 1)
  int* tl_maxe = nullptr;        // entries per row of each tile (longest row, rounded up to 4)
  int2* tl_ent32 = nullptr;      // FP32 entries: (slot of its column, value bits)
  int* tl_eslot = nullptr;       // FP64 entries: slot of its column
  double* tl_ev64 = nullptr;     //               and value
  int64_t* tl_lptr = nullptr;    // X rows each tile loads (ntiles + 1)
  int* tl_lpack = nullptr;       // per X row load: (LDS slot << 24) | row (rows below 2^24)
  int* tl_sslot = nullptr;       // slot of each tile row's own X row (-1: not in its tile)
  int tl_self = 0;               // every tile row's own X row is in its tile's slots
  int tl_trows = 16;             // rows per tile: 16 (256 threads) or 32 (512 threads)
  // stencil form of the tiled copy (a grid operator whose off-diagonal
  // coefficients are one constant per offset): no entry lists, only the
  // diagonal and the offsets' coefficients; X planes of 8 x 8 patches (with
  // their halo) staged in LDS (fe_spmm_tile.hip)
  int st_on = 0;                 // the tiled copy is the stencil form
  int st_p27 = 1;                // 27-point (else 7-point: faces and centre)
  int st_group = 1 << 30;        // pencils per work group (more than the pencils: slice-major)
  int st_nx = 0, st_ny = 0, st_nz = 0;
  double st_w[27] = {};          // coefficient of offset (dz + 1) * 9 + (dy + 1) * 3 + dx + 1 (centre: diagonal)
  double* st_d64 = nullptr;      // diagonal (zero where a row stores none)
  float* st_d32 = nullptr;
  void* st_junk = nullptr;       // 2 x 512 x 16 bytes that rows outside the grid write (never read)
  int use_tl = 0;
};

namespace fe {

inline int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

#define FE_HIP_TRY(expr)                                                        \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess)                                                       \
      return ::fe::fail(-3, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define FE_CHECK_LAUNCH()                                                        \
  do {                                                                           \
    hipError_t _e = hipGetLastError();                                           \
    if (_e != hipSuccess)                                                        \
      return ::fe::fail(-3, std::string("kernel launch: ") + hipGetErrorString(_e)); \
  } while (0)

#define FE_ARG(cond, msg)                              \
  do {                                                 \
    if (!(cond)) return ::fe::fail(-2, msg);           \
  } while (0)

// Line and plane lengths of a 3D grid stencil in natural order (rows
// z * plane + y * line + x), from the column offsets of a sample of rows;
// false for any other structure.
bool detect_grid(const std::vector<int>& rp, const std::vector<int>& ci, int n, int64_t& line, int64_t& plane);

// Ensure scratch slot `slot` of ctx holds at least `bytes`.
int scratch(fe_ctx ctx, int slot, size_t bytes, void** out);

// out[j] = sqrt(sum over c < chunks of part[c * k + j]) for j < k, the chunks
// summed in a fixed order (launched on the context stream).
int reduce_cols_sqrt(fe_ctx ctx, int chunks, int k, const double* part, double* out);

// Looks up (once per device, with ctx->device current) the device address of
// the fixed scratch array behind fe_gram_static and stores it in
// ctx->gram_static. Called by fe_ctx_create.
int gram_static_init(fe_ctx ctx);

// Logical tile index for an XCD-aware launch: the hardware hands workgroups
// to the 8 XCDs round-robin (workgroup b runs on XCD b % 8), so tile
// (b % 8) * per_xcd + b / 8 gives each XCD one contiguous range of tiles and
// keeps that range's neighbours in the XCD's own L2.
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t per_xcd) {
  return (b & 7) * per_xcd + (b >> 3);
}

template <typename T>
struct prec_of;
template <>
struct prec_of<double> {
  static constexpr int value = FE_F64;
};
template <>
struct prec_of<float> {
  static constexpr int value = FE_F32;
};

inline size_t prec_bytes(int prec) { return prec == FE_F64 ? 8 : 4; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// 1 / sqrt(d) for finite d > 0: the hardware estimate and two Newton steps
// y <- y (1 + (1 - d y^2) / 2) (each squares the relative error; the result
// is within a few units in the last place).
__device__ __forceinline__ double rsqrt_nr(double d) {
  double y = __builtin_amdgcn_rsq(d);
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const double e = fma(-d * y, y, 1.0);
    y = fma(0.5 * y, e, y);
  }
  return y;
}

// 1 / d for d whose reciprocal is a finite normal number: the hardware
// estimate and two Newton steps y <- y + y (1 - d y).
__device__ __forceinline__ double rcp_nr(double d) {
  double y = __builtin_amdgcn_rcp(d);
#pragma unroll
  for (int it = 0; it < 2; ++it) y = fma(y, fma(-d, y, 1.0), y);
  return y;
}

inline unsigned blocks_for(int64_t work, int per_block) {
  return static_cast<unsigned>((work + per_block - 1) / per_block);
}

}  // namespace fe
