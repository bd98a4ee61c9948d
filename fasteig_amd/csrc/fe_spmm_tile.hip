// Tiled SpMM: Y = alpha A X + beta X + gamma Z with the X rows of a tile of
// rows staged in LDS.
//
// The rows are grouped in tiles of 16 and the tiles in pencils (runs of
// consecutive tiles). A workgroup takes one pencil and one column slice of
// 256 bytes (64 FP32 or 32 FP64 columns) and walks the pencil's tiles in
// order. LDS holds one 256-byte slot per X row the current tile needs (the
// union of its rows' columns); a slot stays filled while the next tiles still
// need its row, so a tile loads only the rows it adds. The slot plan (which
// row goes to which slot, tile by tile) is computed on the host when the copy
// is built, so the kernel follows lists: the X rows to load for each tile and,
// per nonzero, the slot of its column.
//
// For a 3D grid stencil the tiles are 4 x 4 patches of a plane and a pencil
// runs through consecutive planes: a 27-point tile needs 108 X rows and the
// next plane's tile adds 36 of them, 2.25 row loads per output row against
// 13.5 gathers per row for the 4-row block product and 27 for CSR. The X
// traffic through the vector memory path falls by that factor; the products
// read X from LDS instead. Other matrices take tiles of 16 consecutive rows
// and pencils of 24 tiles, and the copy is built only when that saves loads.
//
// Each row's sum runs over its CSR entries in CSR order, as the CSR kernel's
// does (bitwise the same on a matrix with sorted rows). A tile row shorter
// than the tile's longest row pads with a reserved slot that holds zeros.
#include "fe_common.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <type_traits>

namespace fe {

namespace {

typedef float nt_f4 __attribute__((ext_vector_type(4)));  // 16 bytes for non-temporal access

// Rows per tile: 16 (4 wavefronts x 4 rows of 16 lanes; on a grid a 4 x 4
// patch) or 32 (8 wavefronts; on a grid an 8 x 4 patch). A wider patch loads
// fewer halo rows per output row (27-point stencil: 60 new rows per 32 rows
// against 36 per 16) and shares them inside one workgroup rather than
// through L2 between workgroups that happen to run the neighbouring patch at
// the same plane.
constexpr int TL_ROWS_MAX = 32;
// 16-byte X loads per thread held in registers per tile in flight: at least
// a tile's new rows (36 of a 4 x 4 27-point patch against 3 x 16 rows, 60 of
// an 8 x 4 patch against 2 x 32)
template <int TR>
constexpr int tl_pf() { return TR == 16 ? 3 : 2; }
// minimum waves per SIMD of the FP32 kernel: the LDS of five 16-row or three
// 32-row workgroups fits a CU
template <int TR>
constexpr int tl_waves() { return TR == 16 ? 5 : 6; }
// Tiles per pencil: 24 by default (through 24 planes of a grid; 16 planes ran
// 1-2% slower, 8 planes 6% slower with 16-row tiles), at most tl_pencil_max
// (the LDS tables of a pencil); FE_TILE_PENCIL overrides the default.
constexpr int TL_PENCIL = 24;
template <int TR>
constexpr int tl_pencil_max() { return TR == 16 ? 24 : 32; }
// Edge, in grid points, of the square blocks of patches whose pencils are
// taken together (FE_TILE_BLOCK overrides): 16, i.e. 4 x 4 patches of 4 x 4
// or 2 x 4 patches of 8 x 4.
constexpr int TL_BLOCK_POINTS = 16;
// Pencils per work group: the work items of a group of consecutive pencils
// are taken slice by slice, so the same slice of neighbouring pencils (which
// share halo rows) starts at nearly the same time on one XCD, while the
// group's slices still share its entry lists through L2. C3, 96^3: 4 gave
// 1.7% fewer fetched bytes and 2-4% shorter launches than 1 (all slices of a
// pencil first); 2 and 16 were within 1% of 4.
constexpr int TL_GROUP = 4;
constexpr int TL_MAX_SLOTS = 240;  // LDS slots: with the zero slot and the pencil offsets within 64 KB
static_assert(TL_MAX_SLOTS < 256, "own-row slots are staged as bytes");

// Lane j of each 16-lane row of the wavefront, broadcast to the row (DPP
// row_newbcast).
template <int J>
__device__ __forceinline__ int row_bcast(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x150 + J, 0xF, 0xF, true);
}

// Element (row, col) of a row-major block: with O32 the byte offset fits in
// 32 bits (the host checks the block's extent), which keeps the address
// arithmetic 32-bit on top of the base pointer.
template <bool O32, typename P>
__device__ __forceinline__ P* elem(P* base, int row, int64_t ld, int col) {
  using B = std::conditional_t<std::is_const_v<P>, const char, char>;
  if constexpr (O32)  // rows and ld below 2^24: a full-rate 24-bit multiply
    return reinterpret_cast<P*>(reinterpret_cast<B*>(base) +
                                (__umul24(static_cast<uint32_t>(row), static_cast<uint32_t>(ld)) +
                                 static_cast<uint32_t>(col)) *
                                    static_cast<uint32_t>(sizeof(P)));
  else
    return base + static_cast<int64_t>(row) * ld + col;
}

// base[e] for e >= 0 with the byte offset formed in 32 bits (the base stays
// in scalar registers and the load takes a 32-bit vector offset)
template <typename E, typename B>
__device__ __forceinline__ E ld_off(const B* base, int e) {
  return *reinterpret_cast<const E*>(reinterpret_cast<const char*>(base) +
                                     static_cast<uint32_t>(e) * static_cast<uint32_t>(sizeof(E)));
}

// An entry's slot and value broadcast together: one 64-bit DPP move in FP32
// (slot and value packed), two in FP64.
template <int J>
__device__ __forceinline__ void row_bcast2(int slot, float w, int& s, float& v) {
  const int64_t packed = (static_cast<int64_t>(__float_as_int(w)) << 32) | static_cast<uint32_t>(slot);
  const int64_t b = __builtin_amdgcn_mov_dpp(packed, 0x150 + J, 0xF, 0xF, true);
  s = static_cast<int>(b);
  v = __int_as_float(static_cast<int>(b >> 32));
}
template <int J>
__device__ __forceinline__ void row_bcast2(int slot, double w, int& s, double& v) {
  s = row_bcast<J>(slot);
  v = __longlong_as_double(__builtin_amdgcn_mov_dpp(__double_as_longlong(w), 0x150 + J, 0xF, 0xF, true));
}

template <int J, int END>
struct Unroll {
  template <typename F>
  __device__ __forceinline__ static void run(F& f) {
    f(std::integral_constant<int, J>{});
    Unroll<J + 1, END>::run(f);
  }
};
template <int END>
struct Unroll<END, END> {
  template <typename F>
  __device__ __forceinline__ static void run(F&) {}
};

// Workgroup barrier for the LDS slots only: this wave's LDS operations have
// completed (lgkmcnt 0) when it reaches the barrier, while its global loads
// stay in flight. __syncthreads() also waits for every outstanding global
// load (vmcnt 0), which made each tile of the loop below wait a whole memory
// round trip.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // gfx9 encoding: vmcnt 63, expcnt 7, lgkmcnt 0
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

}  // namespace

// Workgroup: 16 TR threads; thread tid owns tile row tid / 16 and the 16
// bytes at offset 16 * (tid % 16) of the slice. Work items run in groups of
// TL_GROUP pencils, slice by slice, on one XCD: neighbouring pencils share
// halo rows and the group's slices share its entry lists through L2.
//
// Loads are issued ahead of the tile that needs them and no wait covers a
// load issued in the same tile: the new X rows of a tile two tiles ahead
// (held in registers until, after the previous tile's products, the slots
// they go to are free), the entries (two chunks of 16 per row) and the Z and
// X rows of the fused terms one tile ahead, and the load lists three tiles
// ahead. The loop is unrolled by two so that these registers keep fixed roles
// (a register copy of a value still in flight would wait for it), and its
// barriers wait for LDS operations only (with loads issued one tile ahead
// under __syncthreads the kernel took 0.74 of its 0.91 ms with the products
// left out: C3 shape, FP32, 256 columns). A single barrier per tile, with
// the plan keeping a tile's new rows off the previous tile's slots (36 more
// slots for a 27-point stencil), measured 2% slower. The pencil's row, entry
// and load offsets are staged in LDS first.
// FULL: every column slice lies inside the block (k a multiple of the slice
// width), so no load needs a partial-slice path. Every tile's load list holds
// at least the PF x TR rows the registers take (the host pads it with loads
// of a row into a spare slot after the zero slot), so the prefetch loads and
// their LDS writes run without validity checks: a load under a branch, or
// merged with a constant, is waited for where the branches join.
// SELF: each tile row's own X row is among its tile's slots (a stored
// diagonal), so the beta X term reads it from LDS: no second read of X from
// HBM and no registers holding it.
// FZ: the gamma Z term is present. Every iteration of the tile loop issues the
// same loads (the next tile's entries and Z row are fetched also on the last
// tile, from a clamped tile index): with a load under a branch the compiler's
// wait before the first use of a load list covered every load of the previous
// tile (vmcnt(2) instead of vmcnt(7)), so each tile waited for the entries and
// Z rows issued one tile earlier.
// Entries: every row of a tile holds maxe[tile] of them (padded with the zero
// slot), so a lane reads its row's entry e0 + 16 h + sub without a bound check
// or a select (lanes past the row read the next row's entries, which their
// broadcasts never use; the host appends entries after the last tile).
template <typename T, int TR, bool O32, bool FULL, bool SELF, bool FZ>
__global__ __launch_bounds__(16 * TR, (sizeof(T) == 4 && O32 && SELF) ? tl_waves<TR>() : 1) void k_spmm_tile(
    const int64_t* __restrict__ pencil, const int* __restrict__ rows, const int* __restrict__ sslot,
    const int64_t* __restrict__ ebase, const int* __restrict__ maxe, const int2* __restrict__ ent32,
    const int* __restrict__ eslot, const double* __restrict__ ev64,
    const int64_t* __restrict__ lptr, const int* __restrict__ lpack, int zslot, int k,
    int nslices, const T* __restrict__ X, int64_t ldx, T* Y, int64_t ldy, T alpha, T beta, const T* Z, int64_t ldz,
    T gamma, int64_t per_xcd, int64_t nwork) {
  extern __shared__ float4 xs[];  // (zslot + 2) slots of 16 x 16 bytes: zslot holds zeros, zslot + 1 padding loads
  constexpr int PM = tl_pencil_max<TR>();
  __shared__ int s_lp[PM + 1];             // load offsets of the pencil's tiles
  __shared__ int s_eb[PM + 1];             // entry offsets of its tiles
  __shared__ int s_rows[PM * TR];          // its tile rows
  __shared__ int s_maxe[PM];
  // own-row slots of its tile rows (slots < TL_MAX_SLOTS fit a byte: with
  // int entries the workgroup's LDS passed the 32 KB that five per CU allow)
  __shared__ unsigned char s_self[SELF ? PM * TR : 1];
  constexpr int NT = 16 * TR;       // threads
  constexpr int PF = tl_pf<TR>();   // X row loads per thread per tile in registers
  constexpr int V = 16 / (int)sizeof(T);
  constexpr int CW = 16 * V;  // columns per slice
  const int64_t work = xcd_tile(blockIdx.x, per_xcd);
  if (work >= nwork) return;  // the whole workgroup, before any barrier
  // work item -> (pencil, slice): groups of TL_GROUP consecutive pencils,
  // slice-major inside a group (the group's last one may be shorter)
  const int64_t np = nwork / nslices;
  const int64_t gsz = static_cast<int64_t>(TL_GROUP) * nslices;
  const int64_t g = work / gsz;
  const int64_t rem = work - g * gsz;
  const int64_t gp0 = g * TL_GROUP;
  const int64_t gn = min(static_cast<int64_t>(TL_GROUP), np - gp0);
  const int64_t sl = rem / gn;
  const int64_t p = gp0 + (rem - sl * gn);
  const int c0 = static_cast<int>(sl) * CW;
  const int tid = threadIdx.x;
  const int sub = tid & 15;
  const int tr = tid >> 4;
  const int c = c0 + sub * V;  // this thread's first column
  const bool fx = beta != T(0);
  const bool full = FULL || c0 + CW <= k;  // the whole slice lies inside the block
  const int64_t t0 = pencil[p];
  const int nt = static_cast<int>(pencil[p + 1] - t0);  // at most PM
  const int64_t lbase = lptr[t0], eb0 = ebase[t0];
  for (int i = tid; i <= nt; i += NT) s_lp[i] = static_cast<int>(lptr[t0 + i] - lbase);
  for (int i = tid; i <= nt; i += NT) s_eb[i] = static_cast<int>(ebase[t0 + i] - eb0);
  for (int i = tid; i < nt * TR; i += NT) s_rows[i] = rows[t0 * TR + i];
  for (int i = tid; i < nt; i += NT) s_maxe[i] = maxe[t0 + i];
  if (SELF)
    for (int i = tid; i < nt * TR; i += NT) s_self[i] = static_cast<unsigned char>(max(sslot[t0 * TR + i], 0));
  if (tid < 16) xs[zslot * 16 + tid] = make_float4(0.f, 0.f, 0.f, 0.f);
  const int* lp = lpack + lbase;
  auto ld32 = [&](const int* base, int e) { return ld_off<int>(base, e); };
  const int2* e32 = ent32 + eb0;
  const int* es = eslot + eb0;
  const double* ew = ev64 + eb0;
  // entry e of the pencil: slot and value (32-bit byte offsets from the
  // pencil's base: the loads then take the base in SGPRs)
  auto entry = [&](int e, int& slot, T& w) {
    if constexpr (sizeof(T) == 4) {
      const int2 r = ld_off<int2>(e32, e);
      slot = r.x;
      w = __int_as_float(r.y);
    } else {
      slot = ld_off<int>(es, e);
      w = ld_off<double>(ew, e);
    }
  };

  // part `part` (16 bytes) of X row `col` in this slice (zeros past k)
  auto load_x = [&](int col, int part) {
    const int cc = c0 + part * V;
    const T* src = elem<O32>(X, col, ldx, cc);
    float4 f;
    if (full || cc + V <= k) {
      f = *reinterpret_cast<const float4*>(src);
    } else {
      T tmp[V];
#pragma unroll
      for (int q = 0; q < V; ++q) tmp[q] = cc + q < k ? src[q] : T(0);
      __builtin_memcpy(&f, tmp, 16);
    }
    return f;
  };
  // this thread's 16 bytes of row `row` (>= 0) of a block B (zeros outside k)
  auto load_row = [&](const T* B, int64_t ld, int row, T (&v)[V]) {
    if (FULL) {
      const nt_f4 f = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(elem<O32>(B, row, ld, c)));
      __builtin_memcpy(v, &f, 16);
    } else if (c >= k) {
#pragma unroll
      for (int q = 0; q < V; ++q) v[q] = T(0);
    } else if (c + V <= k) {
      const nt_f4 f = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(elem<O32>(B, row, ld, c)));
      __builtin_memcpy(v, &f, 16);
    } else {
#pragma unroll
      for (int q = 0; q < V; ++q) v[q] = c + q < k ? elem<O32>(B, row, ld, c)[q] : T(0);
    }
  };
  // per-tile registers: two chunks of 16 entries of this thread's row (as
  // loaded), and its rows of Z and X
  struct Ahead {
    int ms[2];
    T mw[2];
    T z[V], x[SELF ? 1 : V];
  };
  auto fetch = [&](int tt, Ahead& a) {
    const int e0 = s_eb[tt] + static_cast<int>(__umul24(static_cast<uint32_t>(tr), static_cast<uint32_t>(s_maxe[tt])));
#pragma unroll
    for (int h = 0; h < 2; ++h) entry(e0 + 16 * h + sub, a.ms[h], a.mw[h]);
    const int row = max(s_rows[tt * TR + tr], 0);  // a padding row reads row 0, unused
    if constexpr (FZ) load_row(Z, ldz, row, a.z);
    if constexpr (!SELF) load_row(X, ldx, row, a.x);
  };
  // load list entries (slot << 24 | row) of tile tt for this thread's PF
  // loads, as loaded (a tile past the pencil's last reads the last one's again:
  // its rows then go after the last products, into slots nobody reads)
  auto fetch_list = [&](int tt, int (&pk)[PF]) {
    const int l0 = s_lp[min(tt, nt - 1)] + tr;
#pragma unroll
    for (int j = 0; j < PF; ++j) pk[j] = ld32(lp, l0 + TR * j);
  };
  lds_barrier();
  // the first tile's X rows
  for (int i = tid; i < s_lp[1] * 16; i += NT) {
    const unsigned pk = static_cast<unsigned>(lp[i >> 4]);
    xs[(pk >> 24) * 16 + (i & 15)] = load_x(pk & 0xffffff, i & 15);
  }
  Ahead a[2];                 // entries / Z / X rows of the tiles of each parity
  float4 pre[2][PF];          // new X rows of the tiles of each parity
  int slt[2][PF];             // and their slots (slot * 16 + part)
  // load lists of the tiles two ahead, by the parity of the tile that issues
  // their loads (fixed registers per role: a copy of a register still being
  // loaded would wait for the load)
  int ipk[2][PF];
  // a list's loads: their LDS slots, and the X loads themselves
  auto issue = [&](const int (&pk)[PF], int (&so)[PF], float4 (&po)[PF]) {
#pragma unroll
    for (int j = 0; j < PF; ++j) {
      so[j] = static_cast<int>(static_cast<unsigned>(pk[j]) >> 24) * 16 + sub;
      po[j] = load_x(pk[j] & 0xffffff, sub);
    }
  };
  // the same load order as an iteration of the tile loop (the list of the
  // tile two ahead, then the next tile's entries and Z row), so the waits at
  // the top of the loop hold on the entry path as on the back edge
  {
    int pk[PF];
    fetch_list(1, pk);
    issue(pk, slt[1], pre[1]);
  }
  fetch_list(2, ipk[1]);
  fetch(0, a[0]);
  lds_barrier();
  auto tile = [&](auto Pc, int tt) {
    constexpr int P = decltype(Pc)::value;
    // new X rows of tile tt + 2, the load list of tile tt + 3, and the
    // entries / Z / X rows of tile tt + 1
    issue(ipk[1 - P], slt[P], pre[P]);
    fetch_list(tt + 3, ipk[P]);
    fetch(min(tt + 1, nt - 1), a[1 - P]);  // on the last tile: loads of its own, unused
    // products of tile tt
    const Ahead& cur = a[P];
    const int lim = s_maxe[tt];
    T acc[V];
#pragma unroll
    for (int q = 0; q < V; ++q) acc[q] = T(0);
    for (int ch = 0; ch < lim; ch += 16) {
      int ms;
      T mw;
      if (ch < 32) {
        ms = ch == 0 ? cur.ms[0] : cur.ms[1];
        mw = ch == 0 ? cur.mw[0] : cur.mw[1];
      } else {  // rows longer than 32 entries: later chunks read here
        entry(s_eb[tt] + tr * lim + ch + sub, ms, mw);
      }
      // lim is a multiple of 4 (the host pads the count): whole chunks run
      // without guards, the last one in guarded groups of 4 entries
      auto step = [&](auto J) {
        constexpr int j = decltype(J)::value;
        int s;
        T w;
        row_bcast2<j>(ms, mw, s, w);
        const float4 f = xs[s * 16 + sub];
        T xv[V];
        __builtin_memcpy(xv, &f, 16);
#pragma unroll
        for (int q = 0; q < V; ++q) acc[q] += w * xv[q];
      };
      if (ch + 16 <= lim) {
        Unroll<0, 16>::run(step);
      } else {
        auto group = [&](auto G) {
          constexpr int g = decltype(G)::value;
          if (ch + 4 * g < lim) Unroll<4 * g, 4 * g + 4>::run(step);
        };
        Unroll<0, 4>::run(group);
      }
    }
    const int row = s_rows[tt * TR + tr];
    // the output is formed on every lane (a padding row or a lane past k
    // discards it): the loaded Z / X values are then used on every path, and
    // the next load into their registers needs no wait for them
    T xr[V];  // this row's part of X for the beta term
    if constexpr (SELF) {
      const float4 f = xs[s_self[tt * TR + tr] * 16 + sub];
      __builtin_memcpy(xr, &f, 16);
    } else {
#pragma unroll
      for (int q = 0; q < V; ++q) xr[q] = cur.x[q];
    }
    T out[V];
#pragma unroll
    for (int q = 0; q < V; ++q) {
      out[q] = alpha * acc[q];
      if (fx) out[q] += beta * xr[q];
      if constexpr (FZ) out[q] += gamma * cur.z[q];
    }
    if (row >= 0 && c < k) {
      T* y = elem<O32>(Y, row, ldy, c);
      if (full || c + V <= k) {
        nt_f4 o;
        __builtin_memcpy(&o, out, 16);
        __builtin_nontemporal_store(o, reinterpret_cast<nt_f4*>(y));
      } else {
#pragma unroll
        for (int q = 0; q < V; ++q)
          if (c + q < k) y[q] = out[q];
      }
    }
    lds_barrier();  // every thread is done with tile tt's slots
    // tile tt + 1's new X rows (issued one tile earlier) into their slots
#pragma unroll
    for (int j = 0; j < PF; ++j) xs[slt[1 - P][j]] = pre[1 - P][j];
    if (tt + 1 < nt) {  // rows past the PF registers per thread
      const int l0 = s_lp[tt + 1], nl = s_lp[tt + 2] - l0;
      for (int i = tid + NT * PF; i < nl * 16; i += NT) {
        const unsigned pk = static_cast<unsigned>(lp[l0 + (i >> 4)]);
        xs[(pk >> 24) * 16 + (i & 15)] = load_x(pk & 0xffffff, i & 15);
      }
    }
    lds_barrier();
  };
  for (int tt = 0; tt < nt; tt += 2) {
    tile(std::integral_constant<int, 0>{}, tt);
    if (tt + 1 < nt) tile(std::integral_constant<int, 1>{}, tt + 1);
  }
}

// Stencil form: a grid operator (rows in natural order of an nx x ny x nz
// grid) whose off-diagonal coefficient is one constant per offset within the
// 3 x 3 x 3 neighbourhood, any diagonal, and no entries past the grid's faces.
// No entry lists: the kernel reads the diagonal and takes the coefficients
// from its arguments. A workgroup of 512 threads takes an 8 x 8 patch of a
// plane and one 256-byte column slice (16 lanes x 16 bytes per row; a thread
// computes two neighbouring rows of a line, whose neighbourhoods share two of
// their three columns per line: 4 LDS reads feed both rows' 6 products), and
// walks a pencil of planes. LDS holds three planes of the patch with its halo
// (10 x 10 rows each) as a ring: while plane z is computed, plane z + 1 sits
// in LDS and plane z + 3 is loaded into registers (two planes ahead), so a row
// of X is loaded once per pencil: 100 X rows per 64 output rows reach the
// workgroup, against 60 per 32 for the 8 x 4 entry-list tiles. Halo rows
// outside the grid are zeros in LDS.
//
// Each row's sum runs over the offsets in increasing column order, the order
// of its sorted CSR entries: the product is bitwise the CSR kernel's (a term
// past the grid's faces, or of an offset the row does not store in 27-point
// form, adds coefficient x 0 or 0 x X, which leaves the sum unchanged while
// X is finite).
// Work items: slice-major (every pencil of slice 0, then of slice 1, ...), in
// contiguous ranges per XCD, so an XCD runs one slice of a compact block of
// pencils at a time (64 on C3: five rows of 8 x 8 patches) and their shared
// halo rows come from its L2. C3, FP32 k = 512: 10% fewer fetched bytes and a
// 3% shorter launch than groups of four pencils slice by slice (the order of
// the entry-list tiles, whose slices share entry lists); FP64 k = 512: 13%.
template <typename T>
struct StencilCoef {
  T w[27];
};

constexpr int ST_E = 8;         // patch edge (rows per line and lines per patch)
constexpr int ST_H = ST_E + 2;  // with the halo
constexpr int ST_SLOTS = ST_H * ST_H;
constexpr int ST_PENCIL = 24;   // planes per pencil
// Pencils per work group (FE_STENCIL_GROUP overrides); the default, larger
// than any grid's pencil count, makes the items slice-major.
constexpr int ST_GROUP = 1 << 30;
constexpr int ST_NT = 512;      // threads
constexpr int ST_LD = (ST_SLOTS * 16 + ST_NT - 1) / ST_NT;  // 16-byte X loads per thread per plane (4)

template <typename T, bool O32, bool FULL, bool FZ, bool P27>
__global__ __launch_bounds__(ST_NT, 4) void k_spmm_stencil(
    const T* __restrict__ diag, StencilCoef<T> sc, int nx, int ny, int nz, int npx, int npy, int k, int nslices,
    const T* __restrict__ X, int64_t ldx, T* Y, int64_t ldy, T alpha, T beta, const T* Z, int64_t ldz, T gamma,
    int64_t per_xcd, int64_t nwork, int group, float4* junk) {
  constexpr int V = 16 / (int)sizeof(T);
  constexpr int CW = 16 * V;
  constexpr int PB = (ST_SLOTS + 1) * 16;  // float4 per plane buffer: the slots and a spare one
  __shared__ float4 xs[3 * PB];            // 3 plane buffers (76 KB)
  const int64_t work = xcd_tile(blockIdx.x, per_xcd);
  if (work >= nwork) return;  // the whole workgroup, before any barrier
  // work item -> (pencil, slice): groups of `group` consecutive pencils,
  // slice-major inside a group (the last group may be shorter)
  const int64_t np = nwork / nslices;
  const int64_t gsz = static_cast<int64_t>(group) * nslices;
  const int64_t gi = work / gsz;
  const int64_t grem = work - gi * gsz;
  const int64_t gp0 = gi * group;
  const int64_t gn = min(static_cast<int64_t>(group), np - gp0);
  const int sl = static_cast<int>(grem / gn);
  const int64_t pen = gp0 + (grem - sl * gn);
  const int64_t pp = static_cast<int64_t>(npx) * npy;
  const int zr = static_cast<int>(pen / pp);
  const int rem = static_cast<int>(pen - zr * pp);
  const int py = rem / npx, px = rem - py * npx;
  const int x0 = px * ST_E, y0 = py * ST_E, z0 = zr * ST_PENCIL;
  const int zn = min(nz - z0, ST_PENCIL);  // planes of this pencil
  const int line = nx;
  const int plane = nx * ny;
  const int c0 = sl * CW;
  const int tid = threadIdx.x;
  const int sub = tid & 15;
  const int g = tid >> 4;  // 0..31
  const int c = c0 + sub * V;
  const bool full = FULL || c0 + CW <= k;
  // this thread's halo parts: slots g + 32 j (the last one for g < 4)
  int hoff[ST_LD];
  bool hok[ST_LD];
#pragma unroll
  for (int j = 0; j < ST_LD; ++j) {
    const int s = g + 32 * j;
    const int hy = s / ST_H, hx = s - hy * ST_H;
    const int gx = x0 - 1 + hx, gy = y0 - 1 + hy;
    hok[j] = s < ST_SLOTS && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
    hoff[j] = hok[j] ? gy * line + gx : 0;
  }
  const bool last = g + 32 * (ST_LD - 1) < ST_SLOTS;  // the last part is a slot
  // this thread's two output rows: (xa, y) and (xa + 1, y)
  const int lx = 2 * (g & 3), ly = g >> 2;
  const int xa = x0 + lx, yy = y0 + ly;
  const bool oka = xa < nx && yy < ny, okb = xa + 1 < nx && yy < ny;
  const int roffa = oka ? yy * line + xa : 0;
  const int roffb = okb ? roffa + 1 : roffa;
  const int ctr = (ly + 1) * ST_H + lx + 1;  // slot of row a

  auto load_x = [&](int row) {  // 16 bytes of X row `row` (>= 0) in this slice, zeros past k
    const T* src = elem<O32>(X, row, ldx, c);
    float4 f;
    if (full || c + V <= k) {
      f = *reinterpret_cast<const float4*>(src);
    } else {
      T tmp[V];
#pragma unroll
      for (int q = 0; q < V; ++q) tmp[q] = c + q < k ? src[q] : T(0);
      __builtin_memcpy(&f, tmp, 16);
    }
    return f;
  };
  // plane zz's parts (a plane outside the grid loads a plane inside it and is zeroed when stored)
  // planes past the pencil's last neighbour plane reload that plane (an L2
  // hit): their buffers are never read
  const int zlo = max(z0 - 1, 0), zhi = min(z0 + zn, nz - 1);
  auto load_plane = [&](int zz, float4 (&r)[ST_LD]) {
    const int zc = min(max(zz, zlo), zhi);
#pragma unroll
    for (int j = 0; j < ST_LD; ++j) r[j] = load_x(zc * plane + hoff[j]);
  };
  // every memory operation of the plane loop is unconditional (a store under
  // a branch left the compiler unable to count the loads in flight, and it
  // waited for all of them at the top of each step): a thread without a 4th
  // slot writes the spare slot, rows outside the grid write the junk rows
  const int s3 = last ? g + 32 * (ST_LD - 1) : ST_SLOTS;
  // (selects per component: a select of two float4 objects became a select
  // of their stack addresses)
  auto keep = [](bool ok, const float4& v) {
    return make_float4(ok ? v.x : 0.f, ok ? v.y : 0.f, ok ? v.z : 0.f, ok ? v.w : 0.f);
  };
  auto store_plane = [&](int zz, int buf, const float4 (&r)[ST_LD]) {
    const bool zin = zz >= 0 && zz < nz;
#pragma unroll
    for (int j = 0; j < ST_LD - 1; ++j) xs[buf * PB + (g + 32 * j) * 16 + sub] = keep(zin && hok[j], r[j]);
    xs[buf * PB + s3 * 16 + sub] = keep(zin && hok[ST_LD - 1], r[ST_LD - 1]);
  };
  // Z rows and diagonals of plane zz (clamped into the pencil: past it, unused)
  struct ZD {
    T za[FZ ? V : 1], zb[FZ ? V : 1];
    T da, db;
  };
  auto load_row = [&](int row, T (&v)[FZ ? V : 1]) {
    if constexpr (FZ) {
      if (full || c + V <= k) {
        const nt_f4 f = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(elem<O32>(Z, row, ldz, c)));
        __builtin_memcpy(v, &f, 16);
      } else {
#pragma unroll
        for (int q = 0; q < V; ++q) v[q] = c + q < k ? elem<O32>(Z, row, ldz, c)[q] : T(0);
      }
    }
  };
  auto load_zd = [&](int zz, ZD& o) {
    const int base = min(zz, z0 + zn - 1) * plane;
    o.da = diag[base + roffa];
    o.db = diag[base + roffb];
    load_row(base + roffa, o.za);
    load_row(base + roffb, o.zb);
  };

  {  // planes z0 - 1 and z0 into buffers 0 and 1
    float4 r[ST_LD];
    load_plane(z0 - 1, r);
    store_plane(z0 - 1, 0, r);
    load_plane(z0, r);
    store_plane(z0, 1, r);
  }
  float4 pre[2][ST_LD];  // planes two ahead, by the parity of the step that stores them
  ZD zd[2];
  load_plane(z0 + 1, pre[0]);
  load_plane(z0 + 2, pre[1]);
  load_zd(z0, zd[0]);
  lds_barrier();
  int bp = 0, bc = 1, bn = 2;  // buffers of planes z - 1, z, z + 1

  auto put = [&](T* y, const T (&out)[V]) {
    if (full || c + V <= k) {
      nt_f4 o;
      __builtin_memcpy(&o, out, 16);
      __builtin_nontemporal_store(o, reinterpret_cast<nt_f4*>(y));
    } else {
#pragma unroll
      for (int q = 0; q < V; ++q)
        if (c + q < k) y[q] = out[q];
    }
  };
  auto step = [&](auto Pc, int p) {
    constexpr int P = decltype(Pc)::value;
    const int z = z0 + p;
    store_plane(z + 1, bn, pre[P]);  // its buffer held plane z - 2, free since the last barrier
    lds_barrier();
    load_plane(z + 3, pre[P]);
    load_zd(z + 1, zd[1 - P]);
    T acca[V], accb[V];
#pragma unroll
    for (int q = 0; q < V; ++q) acca[q] = accb[q] = T(0);
    // line l = (dz + 1) * 3 + dy + 1 of the neighbourhood: X at x offsets
    // -1..2 from row a (the dx = -1 and +1 offsets only where the line has
    // them). The reads of line l + 1 are issued before the products of line
    // l, and a barrier that takes the sums after line l's products keeps
    // later reads from moving up: with the products left to sink below all
    // 36 reads the registers passed the 128 that four waves per SIMD allow.
    T xl[2][4][V];
    auto rd = [&](auto L) {
      constexpr int l = decltype(L)::value;
      constexpr int dz = l / 3 - 1, dy = l % 3 - 1;
      constexpr bool side = P27 || (dz == 0 && dy == 0);
      const int buf = dz < 0 ? bp : (dz == 0 ? bc : bn);
      const float4* src = xs + buf * PB + (ctr + dy * ST_H - 1) * 16 + sub;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (side || i == 1 || i == 2) __builtin_memcpy(xl[l & 1][i], &src[16 * i], 16);
    };
    auto fm = [&](auto L) {
      constexpr int l = decltype(L)::value;
      constexpr int dz = l / 3 - 1, dy = l % 3 - 1;
      constexpr bool mid = dz == 0 && dy == 0;
      constexpr bool side = P27 || mid;
      const T(&x)[4][V] = xl[l & 1];
      const T wl = sc.w[l * 3], wc = sc.w[l * 3 + 1], wr = sc.w[l * 3 + 2];
      const T wca = mid ? zd[P].da : wc, wcb = mid ? zd[P].db : wc;
#pragma unroll
      for (int q = 0; q < V; ++q) {
        if constexpr (side) acca[q] += wl * x[0][q];
        acca[q] += wca * x[1][q];
        if constexpr (side) acca[q] += wr * x[2][q];
        if constexpr (side) accb[q] += wl * x[1][q];
        accb[q] += wcb * x[2][q];
        if constexpr (side) accb[q] += wr * x[3][q];
      }
#pragma unroll
      for (int q = 0; q < V; ++q) asm volatile("" : "+v"(acca[q]), "+v"(accb[q])::"memory");
    };
    auto line_terms = [&](auto L) {
      constexpr int l = decltype(L)::value;
      if constexpr (l + 1 < 9) rd(std::integral_constant<int, l + 1>{});
      fm(L);
    };
    rd(std::integral_constant<int, 0>{});
    Unroll<0, 9>::run(line_terms);
    T outa[V], outb[V];
    {
      const float4* xc = xs + bc * PB + ctr * 16 + sub;
      T xra[V], xrb[V];
      __builtin_memcpy(xra, &xc[0], 16);
      __builtin_memcpy(xrb, &xc[16], 16);
#pragma unroll
      for (int q = 0; q < V; ++q) {
        outa[q] = alpha * acca[q];
        outb[q] = alpha * accb[q];
        if (beta != T(0)) {
          outa[q] += beta * xra[q];
          outb[q] += beta * xrb[q];
        }
        if constexpr (FZ) {
          outa[q] += gamma * zd[P].za[q];
          outb[q] += gamma * zd[P].zb[q];
        }
      }
    }
    if constexpr (FULL) {  // every lane's 16 bytes lie inside the block
      T* ya = oka ? elem<O32>(Y, z * plane + roffa, ldy, c) : reinterpret_cast<T*>(junk + tid);
      T* yb = okb ? elem<O32>(Y, z * plane + roffb, ldy, c) : reinterpret_cast<T*>(junk + ST_NT + tid);
      put(ya, outa);
      put(yb, outb);
    } else if (c < k) {
      if (oka) put(elem<O32>(Y, z * plane + roffa, ldy, c), outa);
      if (okb) put(elem<O32>(Y, z * plane + roffb, ldy, c), outb);
    }
    lds_barrier();  // every thread is done with plane z - 1's buffer
    const int t = bp;
    bp = bc;
    bc = bn;
    bn = t;
  };
  for (int p = 0; p < zn; p += 2) {
    step(std::integral_constant<int, 0>{}, p);
    if (p + 1 < zn) step(std::integral_constant<int, 1>{}, p + 1);
  }
}

template <typename T>
static int launch_stencil(fe_ctx ctx, fe_csr A, int k, const T* X, int64_t ldx, T* Y, int64_t ldy, T alpha, T beta,
                          const T* Z, int64_t ldz, T gamma) {
  constexpr int CW = 16 * (16 / (int)sizeof(T));
  if (k > CW && k % CW != 0) {
    // whole slices first, then the remaining columns: the partial-slice
    // kernel's element-wise loads sit under conditions, after which the
    // compiler waits for every load in flight (FP64 per-column closing
    // sweep: 0.73 ms per launch on prefixes of ~170 columns)
    const int kf = k - k % CW;
    const int rc = launch_stencil<T>(ctx, A, kf, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma);
    if (rc) return rc;
    return launch_stencil<T>(ctx, A, k - kf, X + kf, ldx, Y + kf, ldy, alpha, beta, Z ? Z + kf : Z, ldz, gamma);
  }
  const int nslices = (k + CW - 1) / CW;
  const int npx = (A->st_nx + ST_E - 1) / ST_E, npy = (A->st_ny + ST_E - 1) / ST_E;
  const int nzr = (A->st_nz + ST_PENCIL - 1) / ST_PENCIL;
  const int64_t nwork = static_cast<int64_t>(npx) * npy * nzr * nslices;
  const int64_t nb8 = (nwork + 7) / 8 * 8;
  const double lim = 4294967295.0 / sizeof(T);
  const bool o32 = static_cast<double>(A->n) * ldx < lim && static_cast<double>(A->n) * ldy < lim &&
                   (!Z || static_cast<double>(A->n) * ldz < lim);
  const bool full = k % CW == 0;
  const bool fz = Z && gamma != T(0);
  StencilCoef<T> sc;
  for (int o = 0; o < 27; ++o) sc.w[o] = static_cast<T>(A->st_w[o]);
  const T* d = sizeof(T) == 8 ? reinterpret_cast<const T*>(A->st_d64) : reinterpret_cast<const T*>(A->st_d32);
  auto pick = [&](auto zterm, auto p27) {
    constexpr bool F = decltype(zterm)::value, Q = decltype(p27)::value;
    return o32 ? (full ? k_spmm_stencil<T, true, true, F, Q> : k_spmm_stencil<T, true, false, F, Q>)
               : (full ? k_spmm_stencil<T, false, true, F, Q> : k_spmm_stencil<T, false, false, F, Q>);
  };
  auto pick_p = [&](auto zterm) {
    return A->st_p27 ? pick(zterm, std::true_type{}) : pick(zterm, std::false_type{});
  };
  auto kern = fz ? pick_p(std::true_type{}) : pick_p(std::false_type{});
  hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(nb8)), dim3(ST_NT), 0, ctx->stream, d, sc, A->st_nx, A->st_ny,
                     A->st_nz, npx, npy, k, nslices, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma, nb8 / 8, nwork,
                     A->st_group, reinterpret_cast<float4*>(A->st_junk));
  FE_CHECK_LAUNCH();
  return 0;
}

template <typename T>
static int launch_tile(fe_ctx ctx, fe_csr A, int k, const T* X, int64_t ldx, T* Y, int64_t ldy, T alpha, T beta,
                       const T* Z, int64_t ldz, T gamma) {
  constexpr int CW = 16 * (16 / (int)sizeof(T));
  if (k > CW && k % CW != 0) {  // whole slices first, then the rest (as launch_stencil)
    const int kf = k - k % CW;
    const int rc = launch_tile<T>(ctx, A, kf, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma);
    if (rc) return rc;
    return launch_tile<T>(ctx, A, k - kf, X + kf, ldx, Y + kf, ldy, alpha, beta, Z ? Z + kf : Z, ldz, gamma);
  }
  const int nslices = (k + CW - 1) / CW;
  const int64_t nwork = A->tl_npencils * nslices;
  const int64_t nb8 = (nwork + 7) / 8 * 8;
  const size_t lds = static_cast<size_t>(A->tl_slots + 2) * 256;  // the slots, the zero slot and the spare slot
  // 32-bit offsets when every block spans less than 4 GiB
  const double lim = 4294967295.0 / sizeof(T);
  const bool o32 = static_cast<double>(A->n) * ldx < lim && static_cast<double>(A->n) * ldy < lim &&
                   (!Z || static_cast<double>(A->n) * ldz < lim);
  const bool full = k % CW == 0;
  const bool fz = Z && gamma != T(0);
  auto pick = [&](auto rows, auto self, auto zterm) {
    constexpr int R = decltype(rows)::value;
    constexpr bool S = decltype(self)::value, F = decltype(zterm)::value;
    return o32 ? (full ? k_spmm_tile<T, R, true, true, S, F> : k_spmm_tile<T, R, true, false, S, F>)
               : (full ? k_spmm_tile<T, R, false, true, S, F> : k_spmm_tile<T, R, false, false, S, F>);
  };
  auto pick_z = [&](auto rows, auto self) {
    return fz ? pick(rows, self, std::true_type{}) : pick(rows, self, std::false_type{});
  };
  auto pick_s = [&](auto rows) {
    return A->tl_self ? pick_z(rows, std::true_type{}) : pick_z(rows, std::false_type{});
  };
  const bool wide = A->tl_trows == 32;
  auto kern = wide ? pick_s(std::integral_constant<int, 32>{}) : pick_s(std::integral_constant<int, 16>{});
  hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(nb8)), dim3(wide ? 512 : 256), lds, ctx->stream, A->tl_pencil,
                     A->tl_rows, A->tl_sslot, A->tl_ebase, A->tl_maxe, A->tl_ent32, A->tl_eslot, A->tl_ev64,
                     A->tl_lptr, A->tl_lpack, A->tl_slots, k, nslices, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma,
                     nb8 / 8, nwork);
  FE_CHECK_LAUNCH();
  return 0;
}

// Called by fe_spmm for 16-byte aligned blocks when the tiled copy is active.
// fe_csr_use selects it only once the copy is complete (tl_npencils > 0).
int spmm_tiled(fe_ctx ctx, fe_csr A, int prec, int k, const void* X, int64_t ldx, void* Y, int64_t ldy,
               double alpha, double beta, const void* Z, int64_t ldz, double gamma) {
  if (A->st_on) {
    if (prec == FE_F64)
      return launch_stencil<double>(ctx, A, k, (const double*)X, ldx, (double*)Y, ldy, alpha, beta,
                                    (const double*)Z, ldz, gamma);
    return launch_stencil<float>(ctx, A, k, (const float*)X, ldx, (float*)Y, ldy, (float)alpha, (float)beta,
                                 (const float*)Z, ldz, (float)gamma);
  }
  if (prec == FE_F64)
    return launch_tile<double>(ctx, A, k, (const double*)X, ldx, (double*)Y, ldy, alpha, beta, (const double*)Z,
                               ldz, gamma);
  return launch_tile<float>(ctx, A, k, (const float*)X, ldx, (float*)Y, ldy, (float)alpha, (float)beta,
                            (const float*)Z, ldz, (float)gamma);
}

}  // namespace fe

using namespace fe;

namespace {

// Host plan of the tiled copy.
struct TilePlan {
  std::vector<int64_t> pencil{0};
  std::vector<int> rows;
  std::vector<int64_t> ebase{0};  // per tile: its first entry
  std::vector<int> maxe, eslot;   // per tile: entries per row; per entry: slot
  std::vector<double> ev;
  std::vector<int64_t> lptr{0};
  std::vector<int> lpack;  // per X row load: (slot << 24) | row
  std::vector<int> sslot;  // per tile row: slot of its own X row in its tile (-1: none)
  int64_t loads = 0;       // X row loads of the plan (without the padding loads)
  int slots = 0;
  bool self = true;        // every row's own X row is in its tile
};

// Tiles of tr rows (16 or 32) in processing order, grouped in pencils; -1
// pads a tile. A grid whose lines and planes split into px x py patches takes
// patches, any other matrix runs of consecutive rows; false for 32-row tiles
// on a grid that does not split into 8 x 4 patches (its 4 x 4 patches are the
// better plan).
bool make_tiles(int n, int64_t line, int64_t plane, int tr, int pz, int block, std::vector<std::vector<int>>& tiles,
                std::vector<int64_t>& pencil) {
  const int px = tr == 32 ? 8 : 4, py = 4;  // patch extent along the line and across lines
  const bool splits = plane > 0 && line % px == 0 && (plane / line) % py == 0 && n % plane == 0;
  if (plane > 0 && !splits && tr == 32) return false;
  if (splits) {
    // px x py patches; a pencil runs through up to pz consecutive planes; the
    // pencils of one range of planes are taken in blocks of block x block
    // points, row of blocks by row, so the pencils an XCD runs at one time
    // cover a compact area whose shared halo rows its L2 serves
    const int64_t nx = line / px, ny = plane / line / py, nz = n / plane;
    const int64_t bxe = std::max(1, block / px), bye = std::max(1, block / py);
    for (int64_t z0 = 0; z0 < nz; z0 += pz)
      for (int64_t by = 0; by < ny; by += bye)
        for (int64_t bx = 0; bx < nx; bx += bxe)
          for (int64_t yy = by; yy < std::min(ny, by + bye); ++yy)
            for (int64_t xx = bx; xx < std::min(nx, bx + bxe); ++xx) {
              for (int64_t z = z0; z < std::min(nz, z0 + pz); ++z) {
                std::vector<int> t;
                for (int dy = 0; dy < py; ++dy)
                  for (int dx = 0; dx < px; ++dx)
                    t.push_back(static_cast<int>(z * plane + (py * yy + dy) * line + px * xx + dx));
                tiles.push_back(std::move(t));
              }
              pencil.push_back(static_cast<int64_t>(tiles.size()));
            }
    return true;
  }
  const int64_t per_pencil = pz;
  for (int64_t r0 = 0; r0 < n; r0 += tr) {
    std::vector<int> t;
    for (int i = 0; i < tr; ++i) t.push_back(r0 + i < n ? static_cast<int>(r0 + i) : -1);
    tiles.push_back(std::move(t));
    if (static_cast<int64_t>(tiles.size()) % per_pencil == 0) pencil.push_back(static_cast<int64_t>(tiles.size()));
  }
  if (pencil.back() != static_cast<int64_t>(tiles.size())) pencil.push_back(static_cast<int64_t>(tiles.size()));
  return true;
}

// Slot plan: per pencil, a tile keeps the slots of the rows it shares with
// the previous tile and fills freed slots with the rows it adds.
bool plan_slots(int n, const std::vector<int>& rp, const std::vector<int>& ci, const std::vector<double>& v,
                const std::vector<std::vector<int>>& tiles, const std::vector<int64_t>& pencil, int min_loads,
                TilePlan& P) {
  std::vector<int> slot_of(static_cast<size_t>(n), -1), stamp(static_cast<size_t>(n), -1);
  std::vector<int> uni, resident, freelist, kept;
  int slots = 0;
  for (const auto& t : tiles) {  // slots = largest tile union
    uni.clear();
    for (int r : t)
      if (r >= 0) uni.insert(uni.end(), ci.begin() + rp[r], ci.begin() + rp[r + 1]);
    if (uni.size() > t.size() * TL_MAX_SLOTS) return false;  // cannot fit: skip the sort
    std::sort(uni.begin(), uni.end());
    slots = std::max<int>(slots, static_cast<int>(std::unique(uni.begin(), uni.end()) - uni.begin()));
    if (slots > TL_MAX_SLOTS) return false;  // first tile that does not fit ends the plan
  }
  P.slots = slots;
  P.pencil = pencil;
  for (size_t pi = 0; pi + 1 < pencil.size(); ++pi) {
    for (int col : resident) slot_of[col] = -1;
    resident.clear();
    freelist.clear();
    for (int s = slots - 1; s >= 0; --s) freelist.push_back(s);
    for (int64_t ti = pencil[pi]; ti < pencil[pi + 1]; ++ti) {
      const auto& t = tiles[static_cast<size_t>(ti)];
      uni.clear();
      for (int r : t)
        if (r >= 0) uni.insert(uni.end(), ci.begin() + rp[r], ci.begin() + rp[r + 1]);
      std::sort(uni.begin(), uni.end());
      uni.erase(std::unique(uni.begin(), uni.end()), uni.end());
      for (int col : uni) stamp[col] = static_cast<int>(ti);
      kept.clear();
      for (int col : resident) {
        if (stamp[col] == static_cast<int>(ti)) {
          kept.push_back(col);
        } else {
          freelist.push_back(slot_of[col]);
          slot_of[col] = -1;
        }
      }
      resident.swap(kept);
      const size_t first = P.lpack.size();
      for (int col : uni) {
        if (slot_of[col] >= 0) continue;
        slot_of[col] = freelist.back();
        freelist.pop_back();
        resident.push_back(col);
        P.lpack.push_back(static_cast<int>((static_cast<unsigned>(slot_of[col]) << 24) | static_cast<unsigned>(col)));
      }
      P.loads += static_cast<int64_t>(P.lpack.size() - first);
      // at least the loads the kernel's registers take per tile (min_loads):
      // the rest load the tile's first column into the spare slot (slots + 1)
      const int pad_row = uni.empty() ? 0 : uni.front();
      while (P.lpack.size() - first < static_cast<size_t>(min_loads))
        P.lpack.push_back(static_cast<int>((static_cast<unsigned>(slots + 1) << 24) | static_cast<unsigned>(pad_row)));
      P.lptr.push_back(static_cast<int64_t>(P.lpack.size()));
      // every row of the tile holds the longest row's count rounded up to 4
      // entries: its own in CSR order, then the zero slot (slot `slots`)
      int me = 0;
      for (int r : t)
        if (r >= 0) me = std::max(me, rp[r + 1] - rp[r]);
      me = (me + 3) / 4 * 4;
      for (int r : t) {
        P.rows.push_back(r);
        P.sslot.push_back(r >= 0 ? slot_of[r] : -1);
        if (r >= 0 && slot_of[r] < 0) P.self = false;
        const int len = r >= 0 ? rp[r + 1] - rp[r] : 0;
        for (int q = 0; q < len; ++q) {
          P.eslot.push_back(slot_of[ci[rp[r] + q]]);
          P.ev.push_back(v[rp[r] + q]);
        }
        for (int q = len; q < me; ++q) {
          P.eslot.push_back(slots);
          P.ev.push_back(0.0);
        }
      }
      P.maxe.push_back(me);
      P.ebase.push_back(static_cast<int64_t>(P.eslot.size()));
    }
  }
  return true;
}

}  // namespace

namespace {

// Stencil form: the rows are the points of an nx x ny x nz grid in natural
// order, row (x, y, z) stores exactly its in-grid neighbours at the offsets
// of a set S within the 3 x 3 x 3 neighbourhood (columns ascending), each
// offset's coefficient is one constant (bitwise), and the diagonal is any
// value or absent. S and the constants come from the point (1, 1, 1).
struct StencilPlan {
  int nx = 0, ny = 0, nz = 0;
  bool p27 = false;
  double w[27] = {};
  std::vector<double> d;
};

bool detect_stencil(int n, int64_t line, int64_t plane, const std::vector<int>& rp, const std::vector<int>& ci,
                    const std::vector<double>& v, StencilPlan& S) {
  if (line <= 0 || plane <= 0 || plane % line != 0 || n % plane != 0) return false;
  const int nx = static_cast<int>(line), ny = static_cast<int>(plane / line), nz = static_cast<int>(n / plane);
  if (nx < 3 || ny < 3 || nz < 3) return false;
  bool have[27] = {};
  double w[27] = {};
  const int64_t i0 = plane + line + 1;
  for (int p = rp[i0]; p < rp[i0 + 1]; ++p) {
    int64_t off = ci[p] - i0;
    const int64_t dz = (off + plane / 2 + plane) / plane - 1;  // nearest plane
    off -= dz * plane;
    const int64_t dy = (off + line / 2 + line) / line - 1;
    const int64_t dx = off - dy * line;
    if (dz < -1 || dz > 1 || dy < -1 || dy > 1 || dx < -1 || dx > 1) return false;
    const int o = static_cast<int>((dz + 1) * 9 + (dy + 1) * 3 + dx + 1);
    have[o] = o != 13;
    if (o != 13) w[o] = v[p];
  }
  bool p27 = false;
  for (int o = 0; o < 27; ++o) {
    const int dz = o / 9 - 1, dy = (o / 3) % 3 - 1, dx = o % 3 - 1;
    if (have[o] && (dz != 0) + (dy != 0) + (dx != 0) > 1) p27 = true;
  }
  std::vector<double> d(static_cast<size_t>(n), 0.0);
  for (int z = 0, i = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x, ++i) {
        int p = rp[i];
        const int end = rp[i + 1];
        for (int o = 0; o < 27; ++o) {
          const int dz = o / 9 - 1, dy = (o / 3) % 3 - 1, dx = o % 3 - 1;
          const int64_t col = i + dz * plane + dy * line + dx;
          if (o == 13) {  // the diagonal: optional
            if (p < end && ci[p] == col) d[static_cast<size_t>(i)] = v[p++];
            continue;
          }
          const bool in = x + dx >= 0 && x + dx < nx && y + dy >= 0 && y + dy < ny && z + dz >= 0 && z + dz < nz;
          if (!have[o] || !in) continue;
          if (p >= end || ci[p] != col || std::memcmp(&v[p], &w[o], sizeof(double)) != 0) return false;
          ++p;
        }
        if (p != end) return false;  // entries the stencil does not hold
      }
  S.nx = nx;
  S.ny = ny;
  S.nz = nz;
  S.p27 = p27;
  std::memcpy(S.w, w, sizeof(w));
  S.d.swap(d);
  return true;
}

// X rows (inside the grid) the stencil kernel loads per pencil, summed: the
// patch with its halo on the pencil's planes and the two around them
int64_t stencil_loads(const StencilPlan& S) {
  int64_t total = 0;
  for (int z0 = 0; z0 < S.nz; z0 += ST_PENCIL) {
    const int zn = std::min(S.nz - z0, ST_PENCIL);
    const int planes = std::min(z0 + zn, S.nz - 1) - std::max(z0 - 1, 0) + 1;
    for (int y0 = 0; y0 < S.ny; y0 += ST_E)
      for (int x0 = 0; x0 < S.nx; x0 += ST_E) {
        const int wx = std::min(x0 + ST_E, S.nx - 1) - std::max(x0 - 1, 0) + 1;
        const int wy = std::min(y0 + ST_E, S.ny - 1) - std::max(y0 - 1, 0) + 1;
        total += static_cast<int64_t>(planes) * wx * wy;
      }
  }
  return total;
}

}  // namespace

extern "C" int fe_csr_build_tiled(fe_ctx ctx, fe_csr A, double* loads_per_row) {
  FE_ARG(ctx && A, "fe_csr_build_tiled: null argument");
  if (A->tl_npencils > 0) {
    if (loads_per_row) *loads_per_row = A->n ? static_cast<double>(A->tl_loads) / A->n : 0.0;
    return 0;
  }
  const int n = A->n;
  if (n == 0) return 1;
  if (n >= (1 << 24)) return 1;  // a load packs its row in 24 bits (with the slot in the top 8)
  std::vector<int> rp(n + 1), ci(A->nnz);
  std::vector<double> v(A->nnz);
  FE_HIP_TRY(hipMemcpyAsync(rp.data(), A->rowptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, ctx->stream));
  if (A->nnz) {
    FE_HIP_TRY(hipMemcpyAsync(ci.data(), A->colind, sizeof(int) * A->nnz, hipMemcpyDeviceToHost, ctx->stream));
    FE_HIP_TRY(hipMemcpyAsync(v.data(), A->vals64, sizeof(double) * A->nnz, hipMemcpyDeviceToHost, ctx->stream));
  }
  FE_HIP_TRY(hipStreamSynchronize(ctx->stream));
  int64_t line = 0, plane = 0;
  if (!detect_grid(rp, ci, n, line, plane)) line = plane = 0;
  // a constant-coefficient grid operator takes the stencil form (no entry
  // lists); FE_TILE_STENCIL=0 keeps the entry-list tiles
  const char* se = std::getenv("FE_TILE_STENCIL");
  StencilPlan S;
  if (plane > 0 && !(se && std::atoi(se) == 0) && detect_stencil(n, line, plane, rp, ci, v, S)) {
    std::vector<float> d32(S.d.size());
    for (size_t i = 0; i < S.d.size(); ++i) d32[i] = static_cast<float>(S.d[i]);
    if (hipMalloc(&A->st_d64, sizeof(double) * S.d.size()) != hipSuccess ||
        hipMalloc(&A->st_d32, sizeof(float) * S.d.size()) != hipSuccess ||
        hipMalloc(&A->st_junk, 2 * ST_NT * 16) != hipSuccess ||
        hipMemcpyAsync(A->st_d64, S.d.data(), sizeof(double) * S.d.size(), hipMemcpyHostToDevice, ctx->stream) !=
            hipSuccess ||
        hipMemcpyAsync(A->st_d32, d32.data(), sizeof(float) * d32.size(), hipMemcpyHostToDevice, ctx->stream) !=
            hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
      hipStreamSynchronize(ctx->stream);
      if (A->st_d64) hipFree(A->st_d64);
      if (A->st_d32) hipFree(A->st_d32);
      if (A->st_junk) hipFree(A->st_junk);
      A->st_d64 = nullptr;
      A->st_d32 = nullptr;
      A->st_junk = nullptr;
      return fail(-3, "fe_csr_build_tiled: upload failed");
    }
    A->st_on = 1;
    A->st_group = ST_GROUP;
    if (const char* e = std::getenv("FE_STENCIL_GROUP")) A->st_group = std::max(1, std::atoi(e));
    A->st_p27 = S.p27 ? 1 : 0;
    A->st_nx = S.nx;
    A->st_ny = S.ny;
    A->st_nz = S.nz;
    std::memcpy(A->st_w, S.w, sizeof(S.w));
    const int64_t npx = (S.nx + ST_E - 1) / ST_E, npy = (S.ny + ST_E - 1) / ST_E;
    A->tl_npencils = npx * npy * ((S.nz + ST_PENCIL - 1) / ST_PENCIL);
    A->tl_loads = stencil_loads(S);
    A->tl_trows = ST_E * ST_E;
    if (loads_per_row) *loads_per_row = static_cast<double>(A->tl_loads) / n;
    return 0;
  }
  // tile height: 32 rows (8 x 4 patches) on a grid whose lines and planes
  // split into them, else 16; FE_TILE_ROWS=16 or 32 overrides (a 32-row plan
  // that does not fit the LDS slots falls back to 16 rows)
  int tr = plane > 0 ? 32 : 16;
  if (const char* e = std::getenv("FE_TILE_ROWS")) {
    const int want = std::atoi(e);
    if (want == 16 || want == 32) tr = want;
  }
  // pencil length and block edge of the grid order: FE_TILE_PENCIL (clamped to
  // what the kernel's tables hold) and FE_TILE_BLOCK (points) override
  int pz_want = TL_PENCIL, block = TL_BLOCK_POINTS;
  if (const char* e = std::getenv("FE_TILE_PENCIL")) pz_want = std::max(1, std::atoi(e));
  if (const char* e = std::getenv("FE_TILE_BLOCK")) block = std::max(1, std::atoi(e));
  std::vector<std::vector<int>> tiles;
  std::vector<int64_t> pencil{0};
  TilePlan P;
  bool planned = false;
  for (; !planned && tr >= 16; tr /= 2) {
    tiles.clear();
    pencil.assign(1, 0);
    P = TilePlan{};
    const int pf = tr == 16 ? tl_pf<16>() : tl_pf<32>();
    const int pz = std::min(pz_want, tr == 16 ? tl_pencil_max<16>() : tl_pencil_max<32>());
    planned = make_tiles(n, line, plane, tr, pz, block, tiles, pencil) &&
              plan_slots(n, rp, ci, v, tiles, pencil, pf * tr, P);
    if (planned) break;
  }
  if (!planned) return 1;
  const double lpr = static_cast<double>(P.loads) / n;
  if (loads_per_row) *loads_per_row = lpr;
  // worth it only when the loads per row fall below half the nonzeros per row
  if (lpr > 0.5 * static_cast<double>(A->nnz) / n) return 1;
  // entries after the last tile's: a lane reads up to 31 entries past the
  // start of its row (and 15 past its end)
  for (int q = 0; q < 32; ++q) {
    P.eslot.push_back(P.slots);
    P.ev.push_back(0.0);
  }
  std::vector<int2> ent32(P.ev.size());
  for (size_t i = 0; i < P.ev.size(); ++i) {
    const float w = static_cast<float>(P.ev[i]);
    int bits;
    std::memcpy(&bits, &w, 4);
    ent32[i] = make_int2(P.eslot[i], bits);
  }
  auto up = [&](auto** dst, const auto& src) -> int {
    using E = typename std::decay_t<decltype(src)>::value_type;
    const size_t bytes = sizeof(E) * std::max<size_t>(src.size(), 1);
    FE_HIP_TRY(hipMalloc(reinterpret_cast<void**>(dst), bytes));
    if (!src.empty())
      FE_HIP_TRY(hipMemcpyAsync(*dst, src.data(), sizeof(E) * src.size(), hipMemcpyHostToDevice, ctx->stream));
    return 0;
  };
  int rc = 0;
  if ((rc = up(&A->tl_pencil, P.pencil)) || (rc = up(&A->tl_rows, P.rows)) || (rc = up(&A->tl_ebase, P.ebase)) ||
      (rc = up(&A->tl_maxe, P.maxe)) || (rc = up(&A->tl_ent32, ent32)) || (rc = up(&A->tl_eslot, P.eslot)) ||
      (rc = up(&A->tl_ev64, P.ev)) || (rc = up(&A->tl_lptr, P.lptr)) || (rc = up(&A->tl_lpack, P.lpack)) ||
      (rc = up(&A->tl_sslot, P.sslot)) ||
      (rc = hipStreamSynchronize(ctx->stream) == hipSuccess ? 0 : -3)) {
    hipStreamSynchronize(ctx->stream);
    void** arrays[] = {(void**)&A->tl_pencil, (void**)&A->tl_rows,  (void**)&A->tl_ebase, (void**)&A->tl_maxe,
                       (void**)&A->tl_ent32,  (void**)&A->tl_eslot, (void**)&A->tl_ev64,  (void**)&A->tl_lptr,
                       (void**)&A->tl_lpack,  (void**)&A->tl_sslot};
    for (void** a : arrays) {
      if (*a) hipFree(*a);
      *a = nullptr;
    }
    return rc < 0 ? rc : fail(-3, "fe_csr_build_tiled: upload failed");
  }
  A->tl_ntiles = static_cast<int64_t>(tiles.size());
  A->tl_npencils = static_cast<int64_t>(P.pencil.size()) - 1;
  A->tl_loads = P.loads;
  A->tl_slots = P.slots;
  A->tl_self = P.self ? 1 : 0;
  A->tl_trows = tr;
  return 0;
}

extern "C" int fe_csr_tiled_kind(fe_csr A) {
  if (!A || A->tl_npencils <= 0) return 0;
  return A->st_on ? ST_E * ST_E : A->tl_trows;
}
