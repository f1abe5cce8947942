// grid.hip -- structured per-level kernels of the elliptic operator (gfx950).
//
// The CG of Driver.helm_solve (ib_driver_np.py:311-410) runs on the internal
// layout: per level, dense x/y/z edge blocks and a cell block, +axis oriented.
// A p = rot(integr(V(interp(curl p)))) (ib_core_np.py:168-175) is applied in
// three passes.  In each pass the constant-coefficient stencil of each level
// (2.5D tiles staged in LDS, the fastest index k on consecutive lanes) writes
// the regular rows, and the composite rows (level interfaces, outer boundary,
// sliced ELL) write the rows the stencil does not describe.  Cell vectors
// are stored interleaved (x, y, z of a cell adjacent): the composite rows gather
// whole cells, one cache line per cell instead of one per component.
//
//   G1  vc_d(c)  = (U_d(lo) + U_d(hi)) / (2h^2)       U = discrete curl of p
//   V   y(c)     = rou x - (hdt/h^3)(a - b),  a = sum_up (h visc)(x_n - x), b = sum_dn (h visc)(x - x_n)
//                  (a face of the coarsest level's box: x_n = 0 and twice the coefficient)
//   G3  w_e      = sum over the edge's 4 faces of +-(h/2)(y(c1) + y(c2))
//
// Inside a PCG iteration the passes also carry the CG's vector work: the V pass forms
// p.A p from the cells (sum vol x.(V x)), so alpha is known before G3, which then applies
// r -= alpha w without storing w; x += alpha p rides on the direction update.  In float64
// G1 and V run as one pass (x stays in LDS).  Where the data dependences allow, a pass's
// tiles and its composite rows share one launch.
#include "common.h"

#include <type_traits>

namespace pyib {
namespace {

struct Lev {
  int nx, ny, nz;
  int64_t cell_off, e0, e1, e2;
  double h;
  double pc;      // prec on every regular G3 row of the level (0: read prec)
  int hl[3], hh[3];   // hole: cells [hl, hh) belong to the finer level
};

struct Levels {
  int n;
  Lev l[PYIB_MAX_LEVELS];
};

// prec_cls: the PCG's prec array is the one the grid's preconditioner classes describe
Levels levels_of(const pyib_grid* g, bool prec_cls = false) {
  Levels L;
  L.n = g->nlevels;
  for (int i = 0; i < g->nlevels; ++i) {
    const pyib_level& s = g->lev[i];
    L.l[i] = Lev{s.nx, s.ny, s.nz, s.cell_off, s.e_off[0], s.e_off[1], s.e_off[2], s.h,
                 prec_cls ? g->prec_level_g3[i] : 0.0,
                 {s.hole_lo[0], s.hole_lo[1], s.hole_lo[2]}, {s.hole_hi[0], s.hole_hi[1], s.hole_hi[2]}};
  }
  return L;
}

// the per-level preconditioner table, passed by value (kernel arguments)
struct PrecTable {
  double v[PYIB_MAX_LEVELS];
};

inline bool uses_prec_classes(const pyib_grid* g, const void* prec) {
  return g->edge_cls != nullptr && prec != nullptr && prec == g->cls_prec;
}

// ---------------------------------------------------------------- G1
// composite rows in sliced ELL, one wavefront per slice of 64 rows (lane = row).  The slice
// index is wave-uniform (scalar loads of width/sptr); each lane first loads up to kSellBatch
// (column, coefficient) pairs, then issues all their gathers, so a row costs about two
// memory round trips per batch instead of two per entry.  Entries are summed in order.
constexpr int kSellBatch = 8;

// first slice of this wave: wave w of composite block blk (wave-uniform, in a scalar register)
__device__ __forceinline__ int64_t block_wave(int64_t blk) {
  return __builtin_amdgcn_readfirstlane((int)(blk * (blockDim.x >> 6) + (threadIdx.x >> 6)));
}

// ---------------------------------------------------------------- G1
// vc_d(row) = sum_j coef_d * p[col]
// The row bodies take the block's index among the composite blocks (blk of nblk), so a
// launch can also run them beside the tiled pass (see the merged kernels below); B is the
// gather batch (entries are summed in order, so B changes no result, only registers).
template <typename T, int B>
__device__ __forceinline__ void sell_g1_rows(int64_t blk, int64_t nblk, int64_t nslice, const int32_t* __restrict__ row,
                                             const int32_t* __restrict__ width, const int64_t* __restrict__ sptr,
                                             const int32_t* __restrict__ col, const T* __restrict__ coef,
                                             const T* __restrict__ p, T* __restrict__ vc) {
  const int lane = threadIdx.x & 63;
  for (int64_t s = block_wave(blk); s < nslice; s += nblk * (blockDim.x >> 6)) {
    const int W = width[s];
    const int64_t b = sptr[s] + lane;
    const int r = row[s * 64 + lane];
    T a0 = 0, a1 = 0, a2 = 0;
    for (int j0 = 0; j0 < W; j0 += B) {
      int cc[B];
      T c0[B], c1[B], c2[B], xv[B];
#pragma unroll
      for (int u = 0; u < B; ++u) {
        if (j0 + u < W) {
          const int64_t e = b + (int64_t)(j0 + u) * 64;
          cc[u] = col[e];
          c0[u] = coef[3 * e]; c1[u] = coef[3 * e + 1]; c2[u] = coef[3 * e + 2];
        }
      }
#pragma unroll
      for (int u = 0; u < B; ++u)
        if (j0 + u < W) xv[u] = p[cc[u]];
#pragma unroll
      for (int u = 0; u < B; ++u)
        if (j0 + u < W) { a0 += c0[u] * xv[u]; a1 += c1[u] * xv[u]; a2 += c2[u] * xv[u]; }
    }
    if (r >= 0) { vc[3 * (int64_t)r] = a0; vc[3 * (int64_t)r + 1] = a1; vc[3 * (int64_t)r + 2] = a2; }
  }
}


// The same rows with each slice spread over a block's four waves: wave w sums the entries
// j = w, w+4, w+8, ... of its lane's row and the four partial sums are added in wave order.
// The G1 composite holds few rows (the coarse cells around each hole) of about 20 entries, so
// one wave per slice leaves the GPU mostly idle behind long gather chains; four waves per
// slice give four times the waves and a quarter of the chain.
template <typename T>
__global__ __launch_bounds__(256) void k_sell_g1_split(int64_t nslice, const int32_t* __restrict__ row,
                                                       const int32_t* __restrict__ width,
                                                       const int64_t* __restrict__ sptr,
                                                       const int32_t* __restrict__ col, const T* __restrict__ coef,
                                                       const T* __restrict__ p, T* __restrict__ vc) {
  constexpr int B = 6, NW = 4;                  // one batch covers rows of up to 24 entries
  __shared__ T part[NW - 1][3][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int64_t s = blockIdx.x; s < nslice; s += gridDim.x) {          // block-uniform
    const int W = width[s];
    const int64_t b = sptr[s] + lane;
    T a0 = 0, a1 = 0, a2 = 0;
    for (int j0 = wv; j0 < W; j0 += NW * B) {
      int cc[B];
      T c0[B], c1[B], c2[B], xv[B];
#pragma unroll
      for (int u = 0; u < B; ++u) {
        if (j0 + NW * u < W) {
          const int64_t e = b + (int64_t)(j0 + NW * u) * 64;
          cc[u] = col[e];
          c0[u] = coef[3 * e]; c1[u] = coef[3 * e + 1]; c2[u] = coef[3 * e + 2];
        }
      }
#pragma unroll
      for (int u = 0; u < B; ++u)
        if (j0 + NW * u < W) xv[u] = p[cc[u]];
#pragma unroll
      for (int u = 0; u < B; ++u)
        if (j0 + NW * u < W) { a0 += c0[u] * xv[u]; a1 += c1[u] * xv[u]; a2 += c2[u] * xv[u]; }
    }
    if (wv > 0) {
      part[wv - 1][0][lane] = a0;
      part[wv - 1][1][lane] = a1;
      part[wv - 1][2][lane] = a2;
    }
    __syncthreads();
    if (wv == 0) {
      const int r = row[s * 64 + lane];
      if (r >= 0) {
        vc[3 * (int64_t)r] = ((a0 + part[0][0][lane]) + part[1][0][lane]) + part[2][0][lane];
        vc[3 * (int64_t)r + 1] = ((a1 + part[0][1][lane]) + part[1][1][lane]) + part[2][1][lane];
        vc[3 * (int64_t)r + 2] = ((a2 + part[0][2][lane]) + part[1][2][lane]) + part[2][2][lane];
      }
    }
    __syncthreads();                                                   // part is reused
  }
}

// ----------------------------------------------------------------- V
// y(row) = rou x(row) - kv sum_j coef x[col].  With partials, block partials of
// sum_row rw(row) x(row).y(row) (rw = the row's cell volume): these rows' share of p.A p,
// since A = rot integr V interp curl with integr = interp^T diag(1/vol_ci) makes
// p.A p = sum_cells vol (interp curl p).(V interp curl p).
template <typename T, int B>
__device__ __forceinline__ void sell_v_rows(int64_t blk, int64_t nblk, int64_t nslice, const int32_t* __restrict__ row,
                                            const int32_t* __restrict__ width, const int64_t* __restrict__ sptr,
                                            const int32_t* __restrict__ col, const T* __restrict__ coef, T rou, T kv,
                                            const T* __restrict__ x, T* __restrict__ y, const double* __restrict__ rw,
                                            double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  const int lane = threadIdx.x & 63;
  double dot = 0.0;
  for (int64_t s = block_wave(blk); s < nslice; s += nblk * (blockDim.x >> 6)) {
    const int W = width[s];
    const int64_t b = sptr[s] + lane;
    const int r = row[s * 64 + lane];
    T a0 = 0, a1 = 0, a2 = 0;
    for (int j0 = 0; j0 < W; j0 += B) {
      int cc[B];
      T cf[B], x0[B], x1[B], x2[B];
#pragma unroll
      for (int u = 0; u < B; ++u) {
        if (j0 + u < W) {
          const int64_t e = b + (int64_t)(j0 + u) * 64;
          cc[u] = col[e];
          cf[u] = coef[e];
        }
      }
#pragma unroll
      for (int u = 0; u < B; ++u)
        if (j0 + u < W) {
          const int64_t o = 3 * (int64_t)cc[u];
          x0[u] = x[o]; x1[u] = x[o + 1]; x2[u] = x[o + 2];
        }
#pragma unroll
      for (int u = 0; u < B; ++u)
        if (j0 + u < W) { a0 += cf[u] * x0[u]; a1 += cf[u] * x1[u]; a2 += cf[u] * x2[u]; }
    }
    if (r >= 0) {
      const int64_t o = 3 * (int64_t)r;
      const T xc0 = x[o], xc1 = x[o + 1], xc2 = x[o + 2];
      const T y0 = rou * xc0 - kv * a0, y1 = rou * xc1 - kv * a1, y2 = rou * xc2 - kv * a2;
      y[o] = y0;
      y[o + 1] = y1;
      y[o + 2] = y2;
      if (partials)
        dot += rw[s * 64 + lane] * ((double)xc0 * (double)y0 + (double)xc1 * (double)y1 + (double)xc2 * (double)y2);
    }
  }
  if (partials) {
    const double rr = block_sum(dot, sh);
    if (threadIdx.x == 0) partials[blk] = rr;
  }
}
template <typename T>
__global__ void k_sell_v(int64_t nslice, int64_t ncell, const int32_t* __restrict__ row,
                         const int32_t* __restrict__ width, const int64_t* __restrict__ sptr,
                         const int32_t* __restrict__ col, const T* __restrict__ coef, T rou, T kv,
                         const T* __restrict__ x, T* __restrict__ y, const double* __restrict__ rw,
                         double* __restrict__ partials) {
  sell_v_rows<T, kSellBatch>(blockIdx.x, gridDim.x, nslice, row, width, sptr, col, coef, rou, kv, x, y, rw, partials);
}

// ---------------------------------------------------------------- G3
// w(row) = sum_j coef . y[col] (Dirichlet rows 0).  With r non-NULL: r(row) -= alpha w(row)
// instead of storing w, and block partials of r.prec.r over these rows (as the tiled G3 pass)
// mask: Dirichlet rows forced to 0 (NULL: the rows hold none); rprec (with rv): prec at each
// lane's row, stored with the slices, instead of the scattered prec[row] (NULL: read prec)
template <typename T, int B>
__device__ __forceinline__ void sell_g3_rows(int64_t blk, int64_t nblk, int64_t nslice, const int32_t* __restrict__ row,
                                             const int32_t* __restrict__ width, const int64_t* __restrict__ sptr,
                                             const int32_t* __restrict__ col, const T* __restrict__ coef,
                                             const uint8_t* __restrict__ mask, const T* __restrict__ y,
                                             T* __restrict__ w, const double* __restrict__ scal, T* __restrict__ rv,
                                             const T* __restrict__ prec, const T* __restrict__ rprec,
                                             double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  const int lane = threadIdx.x & 63;
  const T alf = rv ? (T)scal[S_ALPHA] : T(0);
  double dot = 0.0;
  for (int64_t s = block_wave(blk); s < nslice; s += nblk * (blockDim.x >> 6)) {
    const int W = width[s];
    const int64_t b = sptr[s] + lane;
    const int r = row[s * 64 + lane];
    T acc = 0;
    for (int j0 = 0; j0 < W; j0 += B) {
      int cc[B];
      T c0[B], c1[B], c2[B], y0[B], y1[B], y2[B];
#pragma unroll
      for (int u = 0; u < B; ++u) {
        if (j0 + u < W) {
          const int64_t e = b + (int64_t)(j0 + u) * 64;
          cc[u] = col[e];
          c0[u] = coef[3 * e]; c1[u] = coef[3 * e + 1]; c2[u] = coef[3 * e + 2];
        }
      }
#pragma unroll
      for (int u = 0; u < B; ++u)
        if (j0 + u < W) {
          const int64_t o = 3 * (int64_t)cc[u];
          y0[u] = y[o]; y1[u] = y[o + 1]; y2[u] = y[o + 2];
        }
#pragma unroll
      for (int u = 0; u < B; ++u) {
        if (j0 + u < W) {
          T t = c0[u] * y0[u] + c1[u] * y1[u];
          t = t + c2[u] * y2[u];
          acc += t;
        }
      }
    }
    if (r >= 0) {
      if (mask && mask[r]) acc = T(0);
      if (rv) {
        const T ri = rv[r] - alf * acc;
        rv[r] = ri;
        const T m = rprec ? rprec[s * 64 + lane] : prec[r];
        dot += (double)(ri * m) * (double)ri;
      } else {
        w[r] = acc;
      }
    }
  }
  if (partials) {
    const double tot = block_sum(dot, sh);
    if (threadIdx.x == 0) partials[blk] = tot;
  }
}


// ------------------------------------------------- 2.5D tiled stencil passes
// Each block owns a TJ x TK tile of one level's (j, k) plane and marches over a chunk of
// CI planes in i (grid: x = tiles of the largest level, y = i-chunks, z = level).  The
// planes it needs (with their halo in j and k) are staged once in LDS per step, so every
// operand is read from memory about once instead of once per stencil tap.  Values outside
// the level box are staged as 0 (only rows away from the box edge are regular).  Only the
// regular rows are written; the composite pass writes the others.
// i-planes per block of the G1/V passes (CI) and of the G3 pass (CI3): measured at D32, 16 planes
// suit the G1/V passes and 8 the G3 pass (240 -> 226 us in float64; 32 planes are slower in both)
#ifndef PYIB_TILE_CI
#define PYIB_TILE_CI 16
#endif
#ifndef PYIB_TILE_CI3
#define PYIB_TILE_CI3 8
#endif
#ifndef PYIB_TILE_PL32
#define PYIB_TILE_PL32 4
#endif
#ifndef PYIB_G3_PL64
#define PYIB_G3_PL64 2
#endif
#ifndef PYIB_G3_PL32
#define PYIB_G3_PL32 2
#endif
constexpr int TJ = 8, TK = 32, CI = PYIB_TILE_CI, CI3 = PYIB_TILE_CI3;   // 256 threads: tj = tid / TK, tk = tid % TK

// grid of the tiled passes; *ci = planes per block (given, doubled until the block count fits
// the reduction partials)
inline dim3 tile_grid(const Levels& L, int* ci) {
  int my = 1, mz = 1, mx = 1;
  for (int i = 0; i < L.n; ++i) {
    my = L.l[i].ny > my ? L.l[i].ny : my;
    mz = L.l[i].nz > mz ? L.l[i].nz : mz;
    mx = L.l[i].nx > mx ? L.l[i].nx : mx;
  }
  const unsigned gx = (unsigned)(((my + TJ - 1) / TJ) * ((mz + TK - 1) / TK));
  int c = *ci;
  while ((int64_t)gx * ((mx + c - 1) / c) * L.n > kRedGrid) c *= 2;
  *ci = c;
  return dim3(gx, (unsigned)((mx + c - 1) / c), (unsigned)L.n);
}

struct Tile {
  int j0, k0, i0, i1, tj, tk;
  bool live;
};
// tile of this block for a level whose (j, k) extent is (ny_t, nz_t) and i extent ni
// block ids remapped so that each XCD (workgroups go round-robin over the 8 XCDs) gets one
// contiguous run of tiles: neighbouring tiles share that XCD's L2 for their halo lines
// b = the block's linear dispatch index among the tile grid tg's blocks
__device__ __forceinline__ void xcd_ids_of(int64_t b, const dim3& tg, int& bx, int& by, int& bz) {
  const int64_t gx = tg.x, gy = tg.y, nb = gx * gy * tg.z;
  const int64_t q = nb >> 3, r = nb & 7, x = b & 7, i = b >> 3;
  const int64_t lb = x * q + (x < r ? x : r) + i;
  bx = (int)(lb % gx);
  by = (int)((lb / gx) % gy);
  bz = (int)(lb / (gx * gy));
}
__device__ __forceinline__ void xcd_ids(int& bx, int& by, int& bz) {
  xcd_ids_of(blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z),
             dim3(gridDim.x, gridDim.y, gridDim.z), bx, by, bz);
}

// every cell the tile owns lies in the level's hole: no regular row (G1, V or G3) belongs to
// such a cell (each of its edges touches it, and it is no real cell), so the tile has nothing to
// write and its block only leaves a zero partial
__device__ __forceinline__ bool in_hole(const Lev& v, int j0, int k0, int i0, int i1) {
  const int j1 = min(j0 + TJ, v.ny), k1 = min(k0 + TK, v.nz);
  return i0 >= v.hl[0] && i1 <= v.hh[0] && j0 >= v.hl[1] && j1 <= v.hh[1] && k0 >= v.hl[2] && k1 <= v.hh[2];
}

// number of the faces of cell (i, j, k) on the level box's boundary.  On a box-face V row (the
// coarsest level's cells on the domain boundary) such a face has a zero ghost and AoL = 2h, twice
// the stencil's coefficient; the staged neighbour outside the box is 0, so the face's extra share
// is -kA x of the cell.
template <typename T>
__device__ __forceinline__ T box_faces(const Lev& v, int i, int j, int k) {
  return (T)((i == 0) + (i == v.nx - 1) + (j == 0) + (j == v.ny - 1) + (k == 0) + (k == v.nz - 1));
}

__device__ __forceinline__ Tile tile_of(const Lev& v, int ci, int bx, int by) {
  Tile t;
  const int ntk = (v.nz + TK - 1) / TK, ntj = (v.ny + TJ - 1) / TJ;
  const int b = bx;
  t.j0 = (b / ntk) * TJ;
  t.k0 = (b % ntk) * TK;
  t.i0 = by * ci;
  t.i1 = min(t.i0 + ci, v.nx);
  t.live = b < ntj * ntk && t.i0 < v.nx && !in_hole(v, t.j0, t.k0, t.i0, t.i1);
  t.tj = threadIdx.x / TK;
  t.tk = threadIdx.x % TK;
  return t;
}

// Staging with a one-plane register prefetch: the loads of the plane needed next are
// issued before the current plane is computed, so their latency overlaps the compute.
// Each thread stages entries q = tid + r*256 (r < R) of a plane of `size` entries.
template <typename T, int R>
struct Prefetch {
  T v[R];
};

// s_waitcnt vmcnt(0): wait until every vector memory operation this wave has issued is done,
// leaving the export and LDS/scalar counters alone.  gfx950 (gfx9) encoding: vmcnt in bits
// 3:0 and 15:14 (0 here), expcnt in bits 6:4 (7: no wait), lgkmcnt in bits 11:8 (15: no wait).
__device__ __forceinline__ void wait_vm_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Nontemporal store.  Measured at D32: the float32 G1 pass's x and the float32 V pass's y go out
// nontemporally (each lowered the float32 iteration by about 6 us); in float64 the fused pass's
// stores stay cached (nontemporal y made it 225 -> 242 us), and so do the G3 pass's r loads and
// stores in both precisions (nontemporal made it 263 -> 337 us in float64).
template <typename T>
__device__ __forceinline__ void st_nt(T* p, T v) {
  __builtin_nontemporal_store(v, p);
}

// ---- G1: vc = interp(curl p) on the cells of plane i (reads x-edges of plane i, y/z-edges
// of planes i and i+1)
// Planes per step of the float32 G1 and V tile passes.  A step waits once for the loads issued
// in the step before (stores share the load counter, so the compiler cannot keep two steps in
// flight), so a wider step puts more bytes in flight per wait.  Measured at D32: 4 planes per
// step 88.7 / 108.2 us (G1 / V), 2 planes 89.3 / 109.8, 1 plane 91.0 / 109.4.
template <typename T>
struct TilePlanes {
  static constexpr int value = sizeof(T) == 4 ? PYIB_TILE_PL32 : 1;
};

template <typename T, int PL>
__device__ __forceinline__ void sg1_tile(int64_t lb, const dim3& tg, int64_t ncell, Levels L, int ci, const uint8_t* __restrict__ reg,
                                               const T* __restrict__ p, T* __restrict__ vc) {
  int bx, by, bz;
  xcd_ids_of(lb, tg, bx, by, bz);
  const Lev& v = L.l[bz];
  const Tile t = tile_of(v, ci, bx, by);
  if (!t.live) return;
  constexpr int XJ = TJ + 1, XK = TK + 1;       // x-edges: rows j..j+1, k..k+1
  constexpr int YJ = TJ, YK = TK + 1;           // y-edges: k..k+1
  constexpr int ZJ = TJ + 1, ZK = TK;           // z-edges: j..j+1
  constexpr int R = 2;                          // every staged plane has < 2*256 entries
  constexpr int NS = PL + 1;                    // y/z-edge planes i..i+PL of a step
  __shared__ T sx[PL][XJ * XK];
  __shared__ T sy[NS][YJ * YK];
  __shared__ T sz[NS][ZJ * ZK];
  const int ny1 = v.ny + 1, nz1 = v.nz + 1;
  // 32-bit offsets inside a level's blocks (each has < 2^31 slots) from per-level bases
  const T* __restrict__ pe0 = p + v.e0;
  const T* __restrict__ pe1 = p + v.e1;
  const T* __restrict__ pe2 = p + v.e2;
  const int j = t.j0 + t.tj, k = t.k0 + t.tk;
  const bool mine = j < v.ny && k < v.nz;
  const uint8_t* __restrict__ rl = reg + v.cell_off;
  T* __restrict__ vl = vc + 3 * v.cell_off;
  // One step computes planes i..i+PL-1.  Its loads, issued one step ahead: the x-edges of those
  // planes, the y/z-edges of planes i+1..i+PL and the flags of this thread's cells; planes past
  // the block's chunk are not loaded.
  struct Pre {
    Prefetch<T, R> x[PL], y[PL], z[PL];
    uint8_t g[PL];
  };
  auto fetch = [&](int i, Pre& f) {
#pragma unroll
    for (int l = 0; l < PL; ++l) {
      const int ix = i + l, iy = i + l + 1;     // block-uniform
      const bool okx = ix < t.i1, oky = iy <= t.i1;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int q = threadIdx.x + r * 256;
        {
          const int jj = t.j0 + q / XK, kk = t.k0 + q % XK;
          const bool in = okx && q < XJ * XK && jj < ny1 && kk < nz1;
          f.x[l].v[r] = in ? pe0[(unsigned)((ix * ny1 + jj) * nz1 + kk)] : T(0);
        }
        {
          const int jj = t.j0 + q / YK, kk = t.k0 + q % YK;
          const bool in = oky && q < YJ * YK && jj < v.ny && kk < nz1;
          f.y[l].v[r] = in ? pe1[(unsigned)((iy * v.ny + jj) * nz1 + kk)] : T(0);
        }
        {
          const int jj = t.j0 + q / ZK, kk = t.k0 + q % ZK;
          const bool in = oky && q < ZJ * ZK && jj < ny1 && kk < v.nz;
          f.z[l].v[r] = in ? pe2[(unsigned)((iy * ny1 + jj) * v.nz + kk)] : T(0);
        }
      }
      f.g[l] = (mine && okx) ? rl[(unsigned)((ix * v.ny + j) * v.nz + k)] : (uint8_t)0;
    }
  };
  auto store = [&](int i, const Pre& f) {
#pragma unroll
    for (int l = 0; l < PL; ++l) {
      const int b = (i + l + 1) % NS;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int q = threadIdx.x + r * 256;
        if (q < XJ * XK) sx[l][q] = f.x[l].v[r];
        if (q < YJ * YK) sy[b][q] = f.y[l].v[r];
        if (q < ZJ * ZK) sz[b][q] = f.z[l].v[r];
      }
    }
  };
  // y/z-edges of plane i0 (slot i0 % NS)
  for (int q = threadIdx.x; q < YJ * YK; q += blockDim.x) {
    const int jj = t.j0 + q / YK, kk = t.k0 + q % YK;
    const bool in = t.i0 <= v.nx && jj < v.ny && kk < nz1;
    sy[t.i0 % NS][q] = in ? pe1[(unsigned)((t.i0 * v.ny + jj) * nz1 + kk)] : T(0);
  }
  for (int q = threadIdx.x; q < ZJ * ZK; q += blockDim.x) {
    const int jj = t.j0 + q / ZK, kk = t.k0 + q % ZK;
    const bool in = t.i0 <= v.nx && jj < ny1 && kk < v.nz;
    sz[t.i0 % NS][q] = in ? pe2[(unsigned)((t.i0 * ny1 + jj) * v.nz + kk)] : T(0);
  }
  const T cf = (T)(1.0 / (2.0 * v.h * v.h));
  Pre f;
  fetch(t.i0, f);
  for (int i = t.i0; i < t.i1; i += PL) {
    // Stores share the load counter, so the compiler waits for every load in flight before the
    // step's first use of the buffer; one explicit wait here (vmcnt(0), the others left alone)
    // keeps it from placing that wait after the step's own loads.
    wait_vm_all();
    store(i, f);
    uint8_t g[PL];
#pragma unroll
    for (int l = 0; l < PL; ++l) g[l] = f.g[l];
    // the next step's loads reuse the registers just stored (no other wave reads them), issued
    // before the barrier; the slots written above were last read before the previous barrier
    if (i + PL < t.i1) fetch(i + PL, f);
    __syncthreads();
#pragma unroll
    for (int l = 0; l < PL; ++l) {
      const int ii = i + l;
      if (ii < t.i1 && g[l]) {
        const int a = ii % NS, b = (ii + 1) % NS;   // y/z slots of planes ii, ii+1
        auto PX = [&](int dj, int dk) { return sx[l][(t.tj + dj) * XK + t.tk + dk]; };
        auto PY = [&](int s, int dk) { return sy[s][t.tj * YK + t.tk + dk]; };
        auto PZ = [&](int s, int dj) { return sz[s][(t.tj + dj) * ZK + t.tk]; };
        const T ux0 = (PY(a, 0) - PY(a, 1)) + (PZ(a, 1) - PZ(a, 0));
        const T ux1 = (PY(b, 0) - PY(b, 1)) + (PZ(b, 1) - PZ(b, 0));
        const T uy0 = (PX(0, 1) - PX(0, 0)) - (PZ(b, 0) - PZ(a, 0));
        const T uy1 = (PX(1, 1) - PX(1, 0)) - (PZ(b, 1) - PZ(a, 1));
        const T uz0 = (PY(b, 0) - PY(a, 0)) - (PX(1, 0) - PX(0, 0));
        const T uz1 = (PY(b, 1) - PY(a, 1)) - (PX(1, 1) - PX(0, 1));
        const unsigned c = 3u * (unsigned)((ii * v.ny + j) * v.nz + k);
        st_nt(vl + c, (ux0 + ux1) * cf);
        st_nt(vl + c + 1, (uy0 + uy1) * cf);
        st_nt(vl + c + 2, (uz0 + uz1) * cf);
      }
    }
    __syncthreads();
  }
}

// ---- G1 + V fused: y = V x with x = interp(curl p) formed in LDS, never read back from HBM.
// Each step computes x on plane ip over the tile and a one-cell halo in j and k (10 x 34
// cells) from the x-edges of plane ip and the y/z-edges of planes ip and ip+1, then applies
// V on plane ip-1 from the x planes ip-2 (kept in registers per cell), ip-1 and ip (LDS).
// Cell flags g1v: bit 0 regular G1 row, bit 1 store x, bit 2 regular V row, bit 3 x from HBM.
// x of a cell comes from: the stencil (bit 0), HBM (bit 3: a real cell whose G1 row is
// composite, written by the composite G1 pass), or 0 (the hole, and outside the box).  x of an
// owned regular cell is stored only when bit 1 is set (the composite V rows read it).  The
// flags of a plane are loaded two steps ahead and the bit-3 x values one step ahead, with the
// edges, so no load waits inside a step; the flags go to LDS with x, where the V step reads its
// own cell's.  With partials: block partials of h^3 sum x.y over the regular V cells (their
// share of p.A p).
template <typename T>
__global__ __launch_bounds__(256) void k_sg1v_t(int64_t ncell, Levels L, int ci, const uint8_t* __restrict__ g1v,
                                                T rou, T visc, T hdt,
                                                const T* __restrict__ p, T* __restrict__ x, T* __restrict__ y,
                                                double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  int bx, by, bz;
  xcd_ids(bx, by, bz);
  const Lev& v = L.l[bz];
  const Tile t = tile_of(v, ci, bx, by);
  double dot = 0.0;
  if (t.live) {
    constexpr int RJ = TJ + 2, RK = TK + 2, RS = RJ * RK;   // cells with the halo: j0-1.., k0-1..
    constexpr int XJ = RJ + 1, XK = RK + 1;                 // x-edges: rows j..j+1, k..k+1
    constexpr int YJ = RJ, YK = RK + 1;                     // y-edges: k..k+1
    constexpr int ZJ = RJ + 1, ZK = RK;                     // z-edges: j..j+1
    constexpr int R = 2;                                    // every staged plane has < 2*256 entries
    __shared__ T sx[XJ * XK];
    __shared__ T sy[2][YJ * YK];
    __shared__ T sz[2][ZJ * ZK];
    __shared__ T sc[2][3][RS];
    __shared__ uint8_t sf[2][RS];
    const int ny1 = v.ny + 1, nz1 = v.nz + 1;
    const int jb = t.j0 - 1, kb = t.k0 - 1;
    auto slot = [](int i) { return (i + 2) & 1; };         // i >= -1
    // Each thread's staging entries sit at the same (j, k) in every plane: their in-plane offsets
    // (32-bit: a level block has < 2^31 slots) and in-box bits are set once for the march.
    // Bit 4r + s of okm: entry r of staged array s (0 x-edges, 1 y-edges, 2 z-edges, 3 cells).
    const int PXs = ny1 * nz1, PYs = v.ny * nz1, PZs = ny1 * v.nz, PCs = v.ny * v.nz;   // plane strides
    const T* __restrict__ pe0 = p + v.e0;
    const T* __restrict__ pe1 = p + v.e1;
    const T* __restrict__ pe2 = p + v.e2;
    const uint8_t* __restrict__ gl = g1v + v.cell_off;
    T* __restrict__ xl = x + 3 * v.cell_off;
    int ox[R], oy[R], oz[R], oc[R];
    unsigned okm = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int q = threadIdx.x + r * 256;
      int j = jb + q / XK, k = kb + q % XK;
      okm |= (unsigned)(q < XJ * XK && j >= 0 && j < ny1 && k >= 0 && k < nz1) << (4 * r);
      ox[r] = j * nz1 + k;
      j = jb + q / YK; k = kb + q % YK;
      okm |= (unsigned)(q < YJ * YK && j >= 0 && j < v.ny && k >= 0 && k < nz1) << (4 * r + 1);
      oy[r] = j * nz1 + k;
      j = jb + q / ZK; k = kb + q % ZK;
      okm |= (unsigned)(q < ZJ * ZK && j >= 0 && j < ny1 && k >= 0 && k < v.nz) << (4 * r + 2);
      oz[r] = j * v.nz + k;
      j = jb + q / RK; k = kb + q % RK;
      okm |= (unsigned)(q < RS && j >= 0 && j < v.ny && k >= 0 && k < v.nz) << (4 * r + 3);
      oc[r] = j * v.nz + k;
    }
    // loads of step ip: x-edges of plane ip, y/z-edges of plane ip+1 (fetch_edges), and x of the
    // plane-ip cells whose flag (loaded one step earlier) says it comes from HBM (bit 3: a
    // composite G1 row; fetch_x)
    auto fetch_edges = [&](int ip, Prefetch<T, R>& fx, Prefetch<T, R>& fy, Prefetch<T, R>& fz) {
      const bool i_x = ip >= 0 && ip < v.nx, i_yz = ip + 1 >= 0 && ip + 1 <= v.nx;   // block-uniform
#pragma unroll
      for (int r = 0; r < R; ++r) {
        fx.v[r] = (i_x && (okm >> (4 * r) & 1)) ? pe0[(unsigned)(ip * PXs + ox[r])] : T(0);
        fy.v[r] = (i_yz && (okm >> (4 * r + 1) & 1)) ? pe1[(unsigned)((ip + 1) * PYs + oy[r])] : T(0);
        fz.v[r] = (i_yz && (okm >> (4 * r + 2) & 1)) ? pe2[(unsigned)((ip + 1) * PZs + oz[r])] : T(0);
      }
    };
    auto fetch_x = [&](int ip, const uint8_t* fl, T (*xc)[3]) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (fl[r] & 8) {                // set only inside the level box (flag loads check it)
          const unsigned c = 3u * (unsigned)(ip * PCs + oc[r]);
          xc[r][0] = xl[c];
          xc[r][1] = xl[c + 1];
          xc[r][2] = xl[c + 2];
        }
      }
    };
    // cell flags of plane ip (0 outside the level box)
    auto fetch_flags = [&](int ip, uint8_t* fl) {
      const bool i_c = ip >= 0 && ip < v.nx;
#pragma unroll
      for (int r = 0; r < R; ++r)
        fl[r] = (i_c && (okm >> (4 * r + 3) & 1)) ? gl[(unsigned)(ip * PCs + oc[r])] : (uint8_t)0;
    };
    // y/z-edges of plane i0-1
    {
      const int ip = t.i0 - 1;
      for (int q = threadIdx.x; q < YJ * YK; q += blockDim.x) {
        const int j = jb + q / YK, k = kb + q % YK;
        const bool in = ip >= 0 && ip <= v.nx && j >= 0 && j < v.ny && k >= 0 && k < nz1;
        sy[slot(ip)][q] = in ? p[v.e1 + ((int64_t)ip * v.ny + j) * nz1 + k] : T(0);
      }
      for (int q = threadIdx.x; q < ZJ * ZK; q += blockDim.x) {
        const int j = jb + q / ZK, k = kb + q % ZK;
        const bool in = ip >= 0 && ip <= v.nx && j >= 0 && j < ny1 && k >= 0 && k < v.nz;
        sz[slot(ip)][q] = in ? p[v.e2 + ((int64_t)ip * ny1 + j) * v.nz + k] : T(0);
      }
    }
    Prefetch<T, R> fx, fy, fz;
    T xc[R][3];                       // x of the fetched plane's bit-3 cells
    uint8_t fA[R], fB[R];             // flags of two consecutive planes (the loop alternates them)
    fetch_flags(t.i0 - 1, fA);
    fetch_edges(t.i0 - 1, fx, fy, fz);
    fetch_x(t.i0 - 1, fA, xc);
    fetch_flags(t.i0, fB);
    const T cf = (T)(1.0 / (2.0 * v.h * v.h));
    const T kA = (T)v.h * visc;
    const T hv = hdt * (T)(1.0 / (v.h * v.h * v.h));
    const T hk = hv * kA;                                    // (hdt / h^3) times the face coefficient
    const int j = t.j0 + t.tj, k = t.k0 + t.tk;
    const int qo = (t.tj + 1) * RK + t.tk + 1;              // this thread's cell in the region
    T xm[3] = {T(0), T(0), T(0)};                            // x of plane ip-2 at this cell
    // one step: fp holds the flags of plane ip and takes those of plane ip + 2; fq holds the
    // flags of plane ip + 1.  The two arrays alternate roles from step to step, so no register
    // that a load in flight writes is copied or overwritten.
    auto step = [&](int ip, uint8_t* fp, const uint8_t* fq) {
      const int a = slot(ip), b = slot(ip + 1);
      uint8_t fc[R];
      // Every load of the previous step is used now (edges and x to LDS, flags for this step's
      // loads), so wait for all of them here (vmcnt(0); expcnt and lgkmcnt left alone).  Without
      // it the compiler put a full wait right after this step's first edge loads, which left the
      // prefetch one load deep.
      wait_vm_all();
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int q = threadIdx.x + r * 256;
        if (q < XJ * XK) sx[q] = fx.v[r];
        if (q < YJ * YK) sy[b][q] = fy.v[r];
        if (q < ZJ * ZK) sz[b][q] = fz.v[r];
        fc[r] = fp[r];
      }
      // the next plane's edges go into the registers just stored, before the barrier: no other
      // wave reads them, and the slots written above were last read before the previous barrier
      if (ip + 1 <= t.i1) fetch_edges(ip + 1, fx, fy, fz);
      __syncthreads();
      // x of the bit-3 cells goes to LDS before the registers take the next plane's loads
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int q = threadIdx.x + r * 256;
        if (q < RS && (fc[r] & 8)) {
#pragma unroll
          for (int d = 0; d < 3; ++d) sc[a][d][q] = xc[r][d];
        }
      }
      if (ip + 1 <= t.i1) {                                  // in flight while plane ip is computed
        fetch_x(ip + 1, fq, xc);
        if (ip + 2 <= t.i1) fetch_flags(ip + 2, fp);
      }
#pragma unroll 1
      for (int r = 0; r < R; ++r) {                          // rolled: fewer live registers
        const int q = threadIdx.x + r * 256;
        if (q < RS) {
          const int jj = q / RK, kk = q % RK;
          T x0 = T(0), x1 = T(0), x2 = T(0);
          const uint8_t fr = r ? fc[1] : fc[0];
          if (fr & 1) {
            auto PX = [&](int dj, int dk) { return sx[(jj + dj) * XK + kk + dk]; };
            auto PY = [&](int s, int dk) { return sy[s][jj * YK + kk + dk]; };
            auto PZ = [&](int s, int dj) { return sz[s][(jj + dj) * ZK + kk]; };
            const T ux0 = (PY(a, 0) - PY(a, 1)) + (PZ(a, 1) - PZ(a, 0));
            const T ux1 = (PY(b, 0) - PY(b, 1)) + (PZ(b, 1) - PZ(b, 0));
            const T uy0 = (PX(0, 1) - PX(0, 0)) - (PZ(b, 0) - PZ(a, 0));
            const T uy1 = (PX(1, 1) - PX(1, 0)) - (PZ(b, 1) - PZ(a, 1));
            const T uz0 = (PY(b, 0) - PY(a, 0)) - (PX(1, 0) - PX(0, 0));
            const T uz1 = (PY(b, 1) - PY(a, 1)) - (PX(1, 1) - PX(0, 1));
            x0 = (ux0 + ux1) * cf;
            x1 = (uy0 + uy1) * cf;
            x2 = (uz0 + uz1) * cf;
            const bool own = jj >= 1 && jj <= TJ && kk >= 1 && kk <= TK && ip >= t.i0 && ip < t.i1;
            if ((fr & 2) && own) {
              const int64_t c = v.cell_off + ((int64_t)ip * v.ny + (jb + jj)) * v.nz + (kb + kk);
              x[3 * c] = x0;
              x[3 * c + 1] = x1;
              x[3 * c + 2] = x2;
            }
          }
          if (!(fr & 8)) {                                   // bit 3: x is already in LDS
            sc[a][0][q] = x0;
            sc[a][1][q] = x1;
            sc[a][2][q] = x2;
          }
          sf[a][q] = fr;
        }
      }
      __syncthreads();
      const int i = ip - 1;                                  // V on plane i
      if (i >= t.i0) {
        const int s0 = slot(i);
        const uint8_t fv = sf[s0][qo];
        if (fv & 4) {
          const int64_t c = v.cell_off + ((int64_t)i * v.ny + j) * v.nz + k;
          const T nbox = (fv & 16) ? box_faces<T>(v, i, j, k) : T(0);
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const T xc0 = sc[s0][d][qo];
            // a - b = kA (differences to the upper neighbours - those from the lower ones), the
            // face coefficient kA factored out of the six faces
            T aa = ((sc[a][d][qo] - xc0) + (sc[s0][d][qo + RK] - xc0)) + (sc[s0][d][qo + 1] - xc0);
            if (fv & 16) aa = aa - nbox * xc0;
            const T bb = ((xc0 - xm[d]) + (xc0 - sc[s0][d][qo - RK])) + (xc0 - sc[s0][d][qo - 1]);
            const T yv = rou * xc0 - hk * (aa - bb);
            y[3 * c + d] = yv;
            dot += (double)xc0 * (double)yv;
          }
        }
      }
      if (ip >= t.i0) {
        const int s0 = slot(ip - 1);
#pragma unroll
        for (int d = 0; d < 3; ++d) xm[d] = sc[s0][d][qo];
      }
    };
    for (int ip = t.i0 - 1; ip <= t.i1; ip += 2) {
      step(ip, fA, fB);
      if (ip + 1 <= t.i1) step(ip + 1, fB, fA);
    }
  }
  if (partials) {
    const double r = block_sum(dot, sh);
    if (threadIdx.x == 0)
      partials[((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = r * (v.h * v.h * v.h);
  }
}

// ---- V: y = rou x - (hdt/h^3)(a - b) on the cells of plane i (planes i-1, i, i+1 of x).
// With partials, block partials of h^3 sum x.y over the regular cells (their share of p.A p,
// as in k_sell_v; vol = h^3 on a regular cell)
template <typename T, int PL>
__device__ __forceinline__ void sv_tile(int64_t lb, const dim3& tg, int64_t ncell, Levels L, int ci, const uint8_t* __restrict__ reg,
                                              T rou, T visc, T hdt, const T* __restrict__ x, T* __restrict__ y,
                                              double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  int bx, by, bz;
  xcd_ids_of(lb, tg, bx, by, bz);
  const Lev& v = L.l[bz];
  const Tile t = tile_of(v, ci, bx, by);
  double dot = 0.0;
  if (t.live) {
    constexpr int PJ = TJ + 2, PK = TK + 2, PS = PJ * PK, R = 2;
    constexpr int NS = PL + 2;                    // x planes i-1..i+PL of a step
    __shared__ T sx[NS][3][PS];                   // [plane slot][component][(j, k) with halo]
    // 32-bit offsets inside the level's cell block from per-level bases
    const T* __restrict__ xl = x + 3 * v.cell_off;
    T* __restrict__ yl = y + 3 * v.cell_off;
    const uint8_t* __restrict__ rl = reg + v.cell_off;
    // x of plane i (0 outside the level box and past the block's last plane + 1)
    auto fetch = [&](int i, Prefetch<T, 3 * R>& f) {
      const bool ok = i >= 0 && i < v.nx && i <= t.i1;   // block-uniform
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int q = threadIdx.x + r * 256;
        const int j = t.j0 - 1 + q / PK, k = t.k0 - 1 + q % PK;
        const bool in = ok && q < PS && j >= 0 && j < v.ny && k >= 0 && k < v.nz;
        const unsigned c = 3u * (unsigned)((i * v.ny + j) * v.nz + k);
#pragma unroll
        for (int d = 0; d < 3; ++d) f.v[3 * r + d] = in ? xl[c + d] : T(0);
      }
    };
    auto store = [&](int i, const Prefetch<T, 3 * R>& f) {
      const int slot = (i + NS) % NS;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int q = threadIdx.x + r * 256;
        if (q < PS) {
#pragma unroll
          for (int d = 0; d < 3; ++d) sx[slot][d][q] = f.v[3 * r + d];
        }
      }
    };
    const T kA = (T)v.h * visc;
    const T hv = hdt * (T)(1.0 / (v.h * v.h * v.h));
    const T hk = hv * kA;                          // (hdt / h^3) times the face coefficient
    const int j = t.j0 + t.tj, k = t.k0 + t.tk;
    const bool mine = j < v.ny && k < v.nz;
    const int q = (t.tj + 1) * PK + t.tk + 1;
    // One step applies V on planes i..i+PL-1.  Its loads, issued one step ahead: x of planes
    // i+1..i+PL and the flags of this thread's cell on planes i..i+PL-1.
    struct Pre {
      Prefetch<T, 3 * R> x[PL];
      uint8_t g[PL];
    };
    auto fetch_all = [&](int i, Pre& f) {
#pragma unroll
      for (int l = 0; l < PL; ++l) {
        fetch(i + l + 1, f.x[l]);
        const int ic = i + l;
        f.g[l] = (mine && ic < t.i1) ? rl[(unsigned)((ic * v.ny + j) * v.nz + k)] : (uint8_t)0;
      }
    };
    {
      Prefetch<T, 3 * R> f0;
      fetch(t.i0 - 1, f0);
      store(t.i0 - 1, f0);
      fetch(t.i0, f0);
      store(t.i0, f0);
    }
    Pre f;
    fetch_all(t.i0, f);
    for (int i = t.i0; i < t.i1; i += PL) {
      // one wait for every load in flight (vmcnt(0), the others left alone): stores share the
      // load counter, so the compiler would wait for all of them anyway, and without this it
      // may do so right after the step's own loads
      wait_vm_all();
      uint8_t g[PL];
#pragma unroll
      for (int l = 0; l < PL; ++l) {
        store(i + l + 1, f.x[l]);
        g[l] = f.g[l];
      }
      // the next step's loads reuse the registers just stored, before the barrier; the slots
      // written above were last read before the previous barrier
      if (i + PL < t.i1) fetch_all(i + PL, f);
      __syncthreads();
#pragma unroll
      for (int l = 0; l < PL; ++l) {
        const int ii = i + l;
        if (ii < t.i1 && g[l]) {
          const unsigned c = 3u * (unsigned)((ii * v.ny + j) * v.nz + k);
          const int sm = (ii - 1 + NS) % NS, s0 = ii % NS, sp = (ii + 1) % NS;
          const T nbox = (g[l] & 2) ? box_faces<T>(v, ii, j, k) : T(0);
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const T xc = sx[s0][d][q];
            // a - b = kA (differences to the upper neighbours - those from the lower ones)
            T a = ((sx[sp][d][q] - xc) + (sx[s0][d][q + PK] - xc)) + (sx[s0][d][q + 1] - xc);
            if (g[l] & 2) a = a - nbox * xc;
            const T b = ((xc - sx[sm][d][q]) + (xc - sx[s0][d][q - PK])) + (xc - sx[s0][d][q - 1]);
            const T yv = rou * xc - hk * (a - b);
            st_nt(yl + c + d, yv);
            dot += (double)xc * (double)yv;
          }
        }
      }
      __syncthreads();
    }
  }
  if (partials) {
    const double r = block_sum(dot, sh);
    if (threadIdx.x == 0)
      partials[lb] = r * (v.h * v.h * v.h);
  }
}

// ---- G3: w on the regular x-, y- and z-edges of plane i (planes i-1 and i of y); the
// composite pass writes the irregular rows; empty slots (and Dirichlet rows when the caller
// leaves them out of the composite) are never written.
// With rr non-NULL the PCG residual update is fused in: w is not stored, r -= alpha w on the
// same rows (alpha = scal[S_ALPHA], already final since p.A p comes from the V pass) and
// block partials of r.prec.r go to partials[block]
// Planes per step of the G3 tile pass (see TilePlanes).  Measured at D32: float64 259.8 us with
// 2 planes, 263.4 with 1; float32 150.4 with 2, 151.6 with 1, 166.5 with 4.
template <typename T>
struct G3Planes {
  static constexpr int value = sizeof(T) == 4 ? PYIB_G3_PL32 : PYIB_G3_PL64;
};

template <typename T, int PL>
__device__ __forceinline__ void sg3_tile(int64_t lb, const dim3& tg, int64_t ncell, Levels L, int ci, const uint8_t* __restrict__ reg,
                                               const T* __restrict__ y, T* __restrict__ w,
                                               const double* __restrict__ scal, T* __restrict__ rr,
                                               const T* __restrict__ prec, double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  int bx, by, bz;
  xcd_ids_of(lb, tg, bx, by, bz);
  const Lev& v = L.l[bz];
  const Tile t = tile_of(v, ci, bx, by);
  double dot = 0.0;
  const T alf = rr ? (T)scal[S_ALPHA] : T(0);
  if (t.live) {
    constexpr int PJ = TJ + 1, PK = TK + 1, PS = PJ * PK, R = 2;   // low-side halo (j-1, k-1)
    constexpr int NS = PL + 1;                    // y planes i-1..i+PL-1 of a step
    __shared__ T sy[NS][3][PS];
    // staged cells keep their (j, k) over the march: 32-bit in-plane offsets and in-box bits
    const int PCs = v.ny * v.nz;
    const T* __restrict__ yl = y + 3 * v.cell_off;
    int oc[R];
    unsigned okm = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int q = threadIdx.x + r * 256;
      const int j = t.j0 - 1 + q / PK, k = t.k0 - 1 + q % PK;
      okm |= (unsigned)(q < PS && j >= 0 && j < v.ny && k >= 0 && k < v.nz) << r;
      oc[r] = j * v.nz + k;
    }
    auto fetch = [&](int i, Prefetch<T, 3 * R>& f) {
      const bool i_in = i >= 0 && i < v.nx;        // block-uniform
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const bool in = i_in && (okm >> r & 1);
        const unsigned c = 3u * (unsigned)(i * PCs + oc[r]);
#pragma unroll
        for (int d = 0; d < 3; ++d) f.v[3 * r + d] = in ? yl[c + d] : T(0);
      }
    };
    auto store = [&](int i, const Prefetch<T, 3 * R>& f) {
      const int slot = (i + NS) % NS;             // i >= -1
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int q = threadIdx.x + r * 256;
        if (q < PS) {
#pragma unroll
          for (int d = 0; d < 3; ++d) sy[slot][d][q] = f.v[3 * r + d];
        }
      }
    };
    const T hh = (T)(0.5 * v.h);
    const int j = t.j0 + t.tj, k = t.k0 + t.tk;
    const bool mine = j < v.ny && k < v.nz;         // this thread's cell is in the level box
    const int q = (t.tj + 1) * PK + t.tk + 1;      // cell (i, j, k) in the staged plane
    const int ny1 = v.ny + 1, nz1 = v.nz + 1;
    // the x-, y- and z-edge (i, j, k) of this thread's cell (32-bit offsets in the level block)
    const int ex = j * nz1 + k, ey = j * nz1 + k, ez = j * v.nz + k;
    const int PXs = ny1 * nz1, PYs = v.ny * nz1, PZs = ny1 * v.nz;
    auto edge = [&](int i, int d) -> int64_t {
      return d == 0 ? v.e0 + (unsigned)(i * PXs + ex)
                    : d == 1 ? v.e1 + (unsigned)(i * PYs + ey) : v.e2 + (unsigned)(i * PZs + ez);
    };
    // per-edge operands of plane i (flag; r and prec in the fused update), loaded one plane ahead.
    // With pc != 0 every regular row of the level has prec == pc, and prec is not read.
    const T pc = (T)v.pc;
    struct EdgeOps {
      T r[3], m[3];
      uint32_t g[3];                     // one register each: packed bytes would make each load wait
    };
    auto efetch = [&](int i, EdgeOps& o) {
      if (!mine) return;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int64_t e = edge(i, d);
        o.g[d] = reg[e];                 // bit 0 tested at use: testing here waits for the load
        if (rr) {
          o.r[d] = rr[e];
          if (pc == T(0)) o.m[d] = prec[e];
        }
      }
    };
    auto emit = [&](const EdgeOps& o, int i, int d, T val) {
      if (!(o.g[d] & 1)) return;
      const int64_t e = edge(i, d);
      if (rr) {
        const T ri = o.r[d] - alf * val;
        rr[e] = ri;
        const T m = pc != T(0) ? pc : o.m[d];
        dot += (double)(ri * m) * (double)ri;
      } else {
        w[e] = val;
      }
    };
    // One step computes the edges of planes i..i+PL-1.  Its loads, issued one step ahead: y of
    // those planes and their edge operands (plane i-1 of y is in LDS from the step before).
    struct Pre {
      Prefetch<T, 3 * R> y[PL];
      EdgeOps e[PL];
    };
    auto fetch_all = [&](int i, Pre& f) {
#pragma unroll
      for (int l = 0; l < PL; ++l) {
        if (i + l < t.i1) {                        // block-uniform
          fetch(i + l, f.y[l]);
          efetch(i + l, f.e[l]);
        }
      }
    };
    {
      Prefetch<T, 3 * R> f0;
      fetch(t.i0 - 1, f0);
      store(t.i0 - 1, f0);
    }
    Pre f{};
    fetch_all(t.i0, f);
    for (int i = t.i0; i < t.i1; i += PL) {
      // every load of the previous step is used now (y to LDS, edge operands): one wait for all
      // of them (vmcnt(0); expcnt and lgkmcnt left alone), and none after this step's loads
      wait_vm_all();
      EdgeOps cur[PL];
#pragma unroll
      for (int l = 0; l < PL; ++l) {
        if (i + l < t.i1) store(i + l, f.y[l]);
        cur[l] = f.e[l];
      }
      // the next step's loads only overwrite this thread's registers, which the lines above have
      // read: they are issued before the barrier
      if (i + PL < t.i1) fetch_all(i + PL, f);
      __syncthreads();
#pragma unroll
      for (int l = 0; l < PL; ++l) {
        const int ii = i + l;
        if (ii < t.i1 && mine) {
          const int s0 = ii % NS, sm = (ii - 1 + NS) % NS;   // planes ii and ii-1
          auto Y = [&](int s, int d, int dj, int dk) { return sy[s][d][q + dj * PK + dk]; };
          {   // x-edge (ii, j, k): -F_y(ii,j,k) + F_y(ii,j,k-1) + F_z(ii,j,k) - F_z(ii,j-1,k)
            const T fy0 = hh * (Y(s0, 1, -1, 0) + Y(s0, 1, 0, 0));
            const T fy1 = hh * (Y(s0, 1, -1, -1) + Y(s0, 1, 0, -1));
            const T fz0 = hh * (Y(s0, 2, 0, -1) + Y(s0, 2, 0, 0));
            const T fz1 = hh * (Y(s0, 2, -1, -1) + Y(s0, 2, -1, 0));
            emit(cur[l], ii, 0, (fy1 - fy0) + (fz0 - fz1));
          }
          {   // y-edge (ii, j, k): F_x(ii,j,k) - F_x(ii,j,k-1) - F_z(ii,j,k) + F_z(ii-1,j,k)
            const T fx0 = hh * (Y(sm, 0, 0, 0) + Y(s0, 0, 0, 0));
            const T fx1 = hh * (Y(sm, 0, 0, -1) + Y(s0, 0, 0, -1));
            const T fz0 = hh * (Y(s0, 2, 0, -1) + Y(s0, 2, 0, 0));
            const T fz1 = hh * (Y(sm, 2, 0, -1) + Y(sm, 2, 0, 0));
            emit(cur[l], ii, 1, (fx0 - fx1) + (fz1 - fz0));
          }
          {   // z-edge (ii, j, k): -F_x(ii,j,k) + F_x(ii,j-1,k) + F_y(ii,j,k) - F_y(ii-1,j,k)
            const T fx0 = hh * (Y(sm, 0, 0, 0) + Y(s0, 0, 0, 0));
            const T fx1 = hh * (Y(sm, 0, -1, 0) + Y(s0, 0, -1, 0));
            const T fy0 = hh * (Y(s0, 1, -1, 0) + Y(s0, 1, 0, 0));
            const T fy1 = hh * (Y(sm, 1, -1, 0) + Y(sm, 1, 0, 0));
            emit(cur[l], ii, 2, (fx1 - fx0) + (fy0 - fy1));
          }
        }
      }
      __syncthreads();
    }
  }
  if (partials) {
    const double s = block_sum(dot, sh);
    if (threadIdx.x == 0) partials[lb] = s;
  }
}

// ---- one launch per pass: the tiled stencil blocks and the composite-row blocks (after the tiles,
// or before them for the float64 G3 pass).  Both write disjoint rows from the same inputs, so the
// composite blocks share the CUs with the tiles instead of waiting for the whole tiled pass.  Each
// tile block keeps its XCD and its partial slot (its index among the tile blocks), and each
// composite block its index among themselves, so every sum is formed in the same order as by two
// launches.  The composite rows gather in batches of
// kMergedBatch, which keeps the merged kernel's registers near the tiled pass's own.
constexpr int kMergedBatch = 4;
// the G3 launch places its composite blocks before the tiles in float64 (measured at D32: 267
// against 278 us) and after them in float32 (163 against 167 us)
template <typename T>
struct G3CompFirst {
  static constexpr bool value = sizeof(T) == 8;
};

template <typename T>
__global__ __launch_bounds__(256) void k_g1_m(int64_t ntile, dim3 tg, int64_t ncell, Levels L, int ci,
                                              const uint8_t* __restrict__ reg, int64_t nslice,
                                              const int32_t* __restrict__ row, const int32_t* __restrict__ width,
                                              const int64_t* __restrict__ sptr, const int32_t* __restrict__ col,
                                              const T* __restrict__ coef, int64_t ncomp, const T* __restrict__ p,
                                              T* __restrict__ vc) {
  if (blockIdx.x < ntile)
    sg1_tile<T, TilePlanes<T>::value>(blockIdx.x, tg, ncell, L, ci, reg, p, vc);
  else
    sell_g1_rows<T, kMergedBatch>(blockIdx.x - ntile, ncomp, nslice, row, width, sptr, col, coef, p, vc);
}

template <typename T>
__global__ __launch_bounds__(256) void k_v_m(int64_t ntile, dim3 tg, int64_t ncell, Levels L, int ci,
                                             const uint8_t* __restrict__ reg, T rou, T visc, T hdt, int64_t nslice,
                                             const int32_t* __restrict__ row, const int32_t* __restrict__ width,
                                             const int64_t* __restrict__ sptr, const int32_t* __restrict__ col,
                                             const T* __restrict__ coef, const double* __restrict__ rw,
                                             int64_t ncomp, const T* __restrict__ x, T* __restrict__ y,
                                             double* __restrict__ tpart, double* __restrict__ cpart) {
  if (blockIdx.x < ntile)
    sv_tile<T, TilePlanes<T>::value>(blockIdx.x, tg, ncell, L, ci, reg, rou, visc, hdt, x, y, tpart);
  else
    sell_v_rows<T, kMergedBatch>(blockIdx.x - ntile, ncomp, nslice, row, width, sptr, col, coef, rou, hdt * visc, x,
                                 y, rw, cpart);
}

template <typename T>
__global__ __launch_bounds__(256) void k_g3_m(int64_t ntile, dim3 tg, int64_t ncell, Levels L, int ci,
                                              const uint8_t* __restrict__ reg, int64_t nslice,
                                              const int32_t* __restrict__ row, const int32_t* __restrict__ width,
                                              const int64_t* __restrict__ sptr, const int32_t* __restrict__ col,
                                              const T* __restrict__ coef, const uint8_t* __restrict__ mask,
                                              int64_t ncomp, const T* __restrict__ y, T* __restrict__ w,
                                              const double* __restrict__ scal, T* __restrict__ rr,
                                              const T* __restrict__ prec, const T* __restrict__ rprec,
                                              double* __restrict__ tpart, double* __restrict__ cpart) {
  if constexpr (G3CompFirst<T>::value) {
    // composite blocks first, padded to a multiple of 8 blocks so each tile block keeps its XCD
    const int64_t nf = (ncomp + 7) & ~int64_t(7);
    if (blockIdx.x >= nf)
      sg3_tile<T, G3Planes<T>::value>(blockIdx.x - nf, tg, ncell, L, ci, reg, y, w, scal, rr, prec, tpart);
    else if (blockIdx.x < ncomp)
      sell_g3_rows<T, kMergedBatch>(blockIdx.x, ncomp, nslice, row, width, sptr, col, coef, mask, y, w, scal, rr,
                                    prec, rprec, cpart);
  } else {
    if (blockIdx.x < ntile)
      sg3_tile<T, G3Planes<T>::value>(blockIdx.x, tg, ncell, L, ci, reg, y, w, scal, rr, prec, tpart);
    else
      sell_g3_rows<T, kMergedBatch>(blockIdx.x - ntile, ncomp, nslice, row, width, sptr, col, coef, mask, y, w, scal,
                                    rr, prec, rprec, cpart);
  }
}

// sum of two partial sets (nb1 from the composite rows, nb2 from the stencil rows)
__global__ void k_alpha2(const double* __restrict__ pa, int na, const double* __restrict__ pb, int nb,
                         double* __restrict__ scal, int write_alpha, double* __restrict__ dot_out) {
  __shared__ double sh[kFinBlock / 64];
  double acc = 0.0;
  for (int b = threadIdx.x; b < na; b += blockDim.x) acc += pa[b];
  double s1 = block_sum(acc, sh);
  acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += pb[b];
  double s2 = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    const double den = s1 + s2;
    if (dot_out) *dot_out = den;
    if (write_alpha) {
      scal[S_DEN] = den;
      scal[S_ALPHA] = scal[S_ETA] / (den != 0.0 ? den : __builtin_inf());
    }
  }
}

// ------------------------------------------------------ CG on the grid
// eta_new = r.prec.r from the two partial rows of the fused G3 passes, beta = eta_new / eta; with
// shift, also eta <- eta_new (the direction update that follows reads only alpha and beta)
__global__ void k_gcg_beta(const double* __restrict__ pa, int na, const double* __restrict__ pb, int nb,
                           double* __restrict__ scal, int shift) {
  __shared__ double sh[kFinBlock / 64];
  double acc = 0.0;
  for (int b = threadIdx.x; b < na; b += blockDim.x) acc += pa[b];
  const double ea = block_sum(acc, sh);
  acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += pb[b];
  const double en = ea + block_sum(acc, sh);
  if (threadIdx.x == 0) {
    scal[S_ETA_NEW] = en;
    const double eta = scal[S_ETA];
    scal[S_BETA] = en / (eta != 0.0 ? eta : __builtin_inf());
    if (shift) scal[S_ETA] = en;
  }
}
template <typename T>
__global__ void k_gcg_dir(int64_t n, const double* __restrict__ scal, const T* __restrict__ r,
                          const T* __restrict__ prec, T* __restrict__ p) {
  const T beta = (T)scal[S_BETA];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = r[i] * prec[i] + beta * p[i];
}
// x += alpha p (alpha of this iteration), then p = r*prec + beta*p
template <typename T>
__global__ void k_gcg_dir_x(int64_t n, const double* __restrict__ scal, const T* __restrict__ r,
                            const T* __restrict__ prec, T* __restrict__ p, T* __restrict__ x) {
  const T alf = (T)scal[S_ALPHA], beta = (T)scal[S_BETA];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T pi = p[i];
    x[i] = x[i] + alf * pi;
    p[i] = r[i] * prec[i] + beta * pi;
  }
}
// the same over the live slots only (cls bit 1), with prec from the level table where cls bit 2
// says it holds: the empty slots and Dirichlet rows keep p = 0 and x as they are.  A float
// thread takes a 16-byte group of 4 slots (one 4-byte class load, then one 16-byte load per
// vector); a double thread takes one slot (measured at D32: 312 us against 350 us for 2-slot
// groups, both over the same bytes).  A group of dead slots is not touched at all.  The class,
// r, p and x streams are loaded and stored nontemporally: each is touched once per iteration
// here, and streaming them past the caches took this kernel from 332 to 285 us at D32 (float64)
// and sped up the passes that follow it, which no longer meet its dirty lines (the float64
// iteration went from 928 to 849 us, the float32 one from 564 to 520 us).
template <typename T>
struct DirGroup {
  static constexpr int W = sizeof(T) == 4 ? 4 : 1;
  using Cls = typename std::conditional<W == 4, uint32_t, uint8_t>::type;
  struct alignas(W * sizeof(T)) V {
    T v[W];
  };
  static constexpr Cls kLive = W == 4 ? (Cls)0x02020202u : (Cls)0x02u;
  typedef float f4 __attribute__((ext_vector_type(4)));
  using NV = typename std::conditional<W == 4, f4, T>::type;   // one load / store per group
  static __device__ __forceinline__ T get(const NV& v, int u) {
    if constexpr (W == 4) return v[u]; else return v;
  }
  static __device__ __forceinline__ void set(NV& v, int u, T a) {
    if constexpr (W == 4) v[u] = a; else v = a;
  }
};
template <typename T>
__device__ __forceinline__ T level_prec(const PrecTable& pt, int c) {
  const int lv = (c >> 4) & 7;
  T m = (T)pt.v[0];
#pragma unroll
  for (int l = 1; l < PYIB_MAX_LEVELS; ++l)
    if (lv == l) m = (T)pt.v[l];
  return m;
}
template <typename T>
__global__ void k_gcg_dir_x_cls(int64_t n, const double* __restrict__ scal, const uint8_t* __restrict__ cls,
                                PrecTable pt, const T* __restrict__ r, const T* __restrict__ prec,
                                T* __restrict__ p, T* __restrict__ x) {
  using VT = DirGroup<T>;
  constexpr int W = VT::W;
  using V = typename VT::V;
  const T alf = (T)scal[S_ALPHA], beta = (T)scal[S_BETA];
  const int64_t nq = n / W, stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (int64_t q = t0; q < nq; q += stride) {
    const uint32_t cw = __builtin_nontemporal_load(reinterpret_cast<const typename VT::Cls*>(cls) + q);
    if (!(cw & VT::kLive)) continue;
    using NV = typename VT::NV;
    const NV pn = __builtin_nontemporal_load(reinterpret_cast<const NV*>(p) + q);
    const NV rn = __builtin_nontemporal_load(reinterpret_cast<const NV*>(r) + q);
    const NV xn = __builtin_nontemporal_load(reinterpret_cast<const NV*>(x) + q);
    V pv, rv, xv;
#pragma unroll
    for (int u = 0; u < W; ++u) { pv.v[u] = VT::get(pn, u); rv.v[u] = VT::get(rn, u); xv.v[u] = VT::get(xn, u); }
    V po = pv;
#pragma unroll
    for (int u = 0; u < W; ++u) {
      const int c = (cw >> (8 * u)) & 0xff;
      if (c & 2) {
        const T m = (c & 4) ? level_prec<T>(pt, c) : prec[W * q + u];
        xv.v[u] = xv.v[u] + alf * pv.v[u];
        po.v[u] = rv.v[u] * m + beta * pv.v[u];
      }
    }
    NV xo, pw;
#pragma unroll
    for (int u = 0; u < W; ++u) { VT::set(xo, u, xv.v[u]); VT::set(pw, u, po.v[u]); }
    __builtin_nontemporal_store(xo, reinterpret_cast<NV*>(x) + q);
    __builtin_nontemporal_store(pw, reinterpret_cast<NV*>(p) + q);
  }
  for (int64_t i = W * nq + t0; i < n; i += stride) {      // the last n % W slots
    const int c = cls[i];
    if (!(c & 2)) continue;
    const T m = (c & 4) ? level_prec<T>(pt, c) : prec[i];
    const T pi = p[i];
    x[i] = x[i] + alf * pi;
    p[i] = r[i] * m + beta * pi;
  }
}
// x += alpha p alone (the last iteration before the host's stop test)
template <typename T>
__global__ void k_gcg_x(int64_t n, const double* __restrict__ scal, const T* __restrict__ p, T* __restrict__ x) {
  const T alf = (T)scal[S_ALPHA];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = x[i] + alf * p[i];
}
__global__ void k_gcg_shift(double* __restrict__ scal) { scal[S_ETA] = scal[S_ETA_NEW]; }
// residual replacement: eta_new <- the replaced residual's eta, beta from it
__global__ void k_gcg_replace(const double* __restrict__ eta_src, double* __restrict__ scal) {
  const double en = eta_src[0];
  const double eta = scal[S_ETA];
  scal[S_ETA_NEW] = en;
  scal[S_BETA] = en / (eta != 0.0 ? eta : __builtin_inf());
}

template <typename T>
__global__ void k_gcg_begin(int64_t n, const T* __restrict__ res, const T* __restrict__ s,
                            const uint8_t* __restrict__ mask, const T* __restrict__ prec, T* __restrict__ r,
                            T* __restrict__ x, T* __restrict__ p, double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  double rr = 0, pr = 0, xx = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T ri = mask[i] ? T(0) : res[i];
    const T xi = s[i];
    const T pi = ri * prec[i];
    r[i] = ri; x[i] = xi; p[i] = pi;
    rr += (double)ri * (double)ri; pr += (double)pi * (double)ri; xx += (double)xi * (double)xi;
  }
  double a = block_sum(rr, sh), b = block_sum(pr, sh), c = block_sum(xx, sh);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = a; partials[kRedGrid + blockIdx.x] = b; partials[2 * kRedGrid + blockIdx.x] = c;
  }
}
__global__ void k_gcg_begin_fin(const double* __restrict__ partials, int nb, double tol1, double tol2,
                                double* __restrict__ scal) {
  __shared__ double sh[kFinBlock / 64];
  double v[3];
  for (int k = 0; k < 3; ++k) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += partials[k * kRedGrid + b];
    v[k] = block_sum(acc, sh);
  }
  if (threadIdx.x == 0) {
    const double rr = v[0], eta = v[1], xnorm = v[2] + 1e-16;
    scal[S_L2I] = rr; scal[S_ETA] = eta;
    scal[S_ERR] = xnorm * (tol1 * tol1) + (tol2 * tol2) * eta;
  }
}
template <typename T>
__global__ void k_gsub(int64_t n, const T* __restrict__ a, const T* __restrict__ b, T* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = a[i] - b[i];
}
template <typename T>
__global__ void k_gtrue_res(int64_t n, const T* __restrict__ res, const T* __restrict__ img,
                            const uint8_t* __restrict__ mask, const T* __restrict__ prec, T* __restrict__ tr,
                            double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  double a = 0, b = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T t = mask[i] ? T(0) : res[i] - img[i];
    tr[i] = t;
    a += (double)(t * prec[i]) * (double)t;
    b += (double)t * (double)t;
  }
  double ra = block_sum(a, sh), rb = block_sum(b, sh);
  if (threadIdx.x == 0) { partials[blockIdx.x] = ra; partials[kRedGrid + blockIdx.x] = rb; }
}

// signed gather between numberings: out[i] = idx[i] >= 0 ? sigma[i] * in[idx[i]] : 0
template <typename A, typename B>
__global__ void k_permute(int64_t n, const int64_t* __restrict__ idx, const int8_t* __restrict__ sigma,
                          const A* __restrict__ in, B* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = idx[i];
    out[i] = s >= 0 ? (B)((A)sigma[i] * in[s]) : B(0);
  }
}

// ------------------------------------------------------- launchers
inline int slice_grid(int64_t nslice) { return grid_for(nslice * 64); }
// one wavefront per slice, capped at the partials a reduction holds (the kernels stride over
// the slices)
inline int slice_red_grid(int64_t nslice) {
  const int g = slice_grid(nslice);
  return g > kRedGrid ? kRedGrid : g;
}
// G1 and V passes of A p: y = V x, x = interp curl p, into ws->cell_b.  In float64 x reaches
// ws->cell_a only on the composite G1 rows and the cells the composite V rows read (fused
// pass); in float32 the two tiled passes go through ws->cell_a.  With want_dot,
// partials of p.A p go to ws->partials [0, *nv) (composite V rows) and
// [kRedGrid, kRedGrid + *nt) (tiled pass): p.A p = sum_cells vol x.(V x), so the dot needs no
// second read of p.  It equals p.w when p is zero on the Dirichlet rows, as every PCG
// direction is.  Tiled and composite kernels write disjoint rows; running them on separate
// streams measured no faster than in order.
template <typename T>
void g1_v(const pyib_grid* g, const pyib_work* ws, double rou, double visc, double hdt, const T* p, bool want_dot,
          int* nv, int* nt, hipStream_t st) {
  const Levels L = levels_of(g);
  T* ca = static_cast<T*>(ws->cell_a);
  T* cb = static_cast<T*>(ws->cell_b);
  int ci = CI;
  const dim3 gt = tile_grid(L, &ci);
  double* tpart = want_dot ? ws->partials + kRedGrid : nullptr;
  double* cpart = want_dot ? ws->partials : nullptr;
  const int64_t ntile = (int64_t)gt.x * gt.y * gt.z;
  const int n1 = slice_red_grid(g->v_nslice);
  const int64_t nc1 = g->v_nslice > 0 ? n1 : 0;
  if constexpr (sizeof(T) == 8) {
    // float64: the composite G1 rows first (the fused pass reads their x from HBM), then one
    // fused G1+V pass, then the composite V rows (they read the x the fused pass stores)
    if (g->g1_nslice > 0)
      k_sell_g1_split<T><<<grid_for(g->g1_nslice * kBlock), kBlock, 0, st>>>(g->g1_nslice, g->g1_row, g->g1_width,
                                                                            g->g1_sptr, g->g1_col,
                                                                            (const T*)g->g1_coef, p, ca);
    k_sg1v_t<T><<<gt, kBlock, 0, st>>>(g->ncell, L, ci, g->cell_g1v, (T)rou, (T)visc, (T)hdt, p, ca, cb, tpart);
    if (nc1 > 0)
      k_sell_v<T><<<n1, kBlock, 0, st>>>(g->v_nslice, g->ncell, g->v_row, g->v_width, g->v_sptr, g->v_col,
                                         (const T*)g->v_coef, (T)rou, (T)(hdt * visc), ca, cb, g->v_rw, cpart);
  } else {
    // float32: G1 (tiles + composite rows) then V (tiles + composite rows), x through HBM; the
    // fused pass is latency-bound once its bytes halve
    const int64_t nc0 = g->g1_nslice > 0 ? slice_grid(g->g1_nslice) : 0;
    k_g1_m<T><<<(unsigned)(ntile + nc0), kBlock, 0, st>>>(ntile, gt, g->ncell, L, ci, g->cell_reg_g1, g->g1_nslice,
                                                         g->g1_row, g->g1_width, g->g1_sptr, g->g1_col,
                                                         (const T*)g->g1_coef, nc0, p, ca);
    k_v_m<T><<<(unsigned)(ntile + nc1), kBlock, 0, st>>>(ntile, gt, g->ncell, L, ci, g->cell_reg, (T)rou, (T)visc,
                                                        (T)hdt, g->v_nslice, g->v_row, g->v_width, g->v_sptr,
                                                        g->v_col, (const T*)g->v_coef, g->v_rw, nc1, ca, cb, tpart,
                                                        cpart);
  }
  *nv = (int)nc1;
  *nt = (int)ntile;
}

// G3 pass from ws->cell_b: w = rot integr y, or with rr non-NULL the fused residual update
// rr -= alpha w with r.prec.r partials in ws->partials [2 kRedGrid, + *nt) (tiled) and
// [3 kRedGrid, + *ns) (composite rows)
template <typename T>
void g3(const pyib_grid* g, const pyib_work* ws, T* w, const double* scal, T* rr, const T* prec, int* nt, int* ns,
        hipStream_t st) {
  const bool cls = rr && uses_prec_classes(g, prec);
  const Levels L = levels_of(g, cls);
  const T* cb = static_cast<const T*>(ws->cell_b);
  int ci = CI3;
  const dim3 gt = tile_grid(L, &ci);
  const int n3 = slice_red_grid(g->g3_nslice);
  const int64_t ntile = (int64_t)gt.x * gt.y * gt.z, nc3 = g->g3_nslice > 0 ? n3 : 0;
  const int64_t nc3_grid = G3CompFirst<T>::value ? ((nc3 + 7) & ~int64_t(7)) : nc3;
  k_g3_m<T><<<(unsigned)(ntile + nc3_grid), kBlock, 0, st>>>(ntile, gt, g->ncell, L, ci, cls ? g->edge_cls : g->edge_reg,
                                                       g->g3_nslice, g->g3_row, g->g3_width, g->g3_sptr, g->g3_col,
                                                       (const T*)g->g3_coef, g->g3_mask, nc3, cb, w, scal, rr,
                                                       prec, cls ? (const T*)g->g3_rprec : nullptr,
                                                       rr ? ws->partials + 2 * kRedGrid : nullptr,
                                                       rr ? ws->partials + 3 * kRedGrid : nullptr);
  if (nt) *nt = (int)ntile;
  if (ns) *ns = (int)nc3;
}

// w = A p on the grid (see g1_v for the dot partials)
template <typename T>
int gflow(const pyib_grid* g, const pyib_work* ws, double rou, double visc, double hdt, const T* p, T* w,
          bool want_dot, int* nb1, int* nb2, hipStream_t st) {
  g1_v<T>(g, ws, rou, visc, hdt, p, want_dot, nb1, nb2, st);
  g3<T>(g, ws, w, nullptr, nullptr, nullptr, nullptr, nullptr, st);
  return 0;
}

// n_iter PCG iterations.  Per iteration: G1, V (+ p.A p partials), alpha, G3 fused with
// r -= alpha w (+ r.prec.r partials), beta, then x += alpha p with p = r prec + beta p in
// one pass (or x alone before the host's stop test when skip_last)
template <typename T>
int gcg_iterate(const pyib_grid* g, const pyib_work* ws, const pyib_cg* cg, double rou, double visc, double hdt,
                int n_iter, int skip_last, hipStream_t st) {
  T* x = static_cast<T*>(cg->x);
  T* r = static_cast<T*>(cg->r);
  T* p = static_cast<T*>(cg->p);
  const T* prec = static_cast<const T*>(cg->prec);
  const int64_t n = g->nedge;
  const bool cls = uses_prec_classes(g, cg->prec);
  PrecTable pt{};
  for (int l = 0; l < PYIB_MAX_LEVELS; ++l) pt.v[l] = g->prec_level[l];
  for (int it = 0; it < n_iter; ++it) {
    int nv, nt, nt3, ns3;
    g1_v<T>(g, ws, rou, visc, hdt, p, true, &nv, &nt, st);
    k_alpha2<<<1, kFinBlock, 0, st>>>(ws->partials, nv, ws->partials + kRedGrid, nt, cg->scal, 1, nullptr);
    g3<T>(g, ws, nullptr, cg->scal, r, prec, &nt3, &ns3, st);
    const bool last = skip_last && it == n_iter - 1;
    k_gcg_beta<<<1, kFinBlock, 0, st>>>(ws->partials + 2 * kRedGrid, nt3, ws->partials + 3 * kRedGrid, ns3,
                                        cg->scal, last ? 0 : 1);
    if (!last) {
      if (cls)
        k_gcg_dir_x_cls<T><<<grid_for((n + DirGroup<T>::W - 1) / DirGroup<T>::W), kBlock, 0, st>>>(
            n, cg->scal, g->edge_cls, pt, r, prec, p, x);
      else
        k_gcg_dir_x<T><<<grid_for(n), kBlock, 0, st>>>(n, cg->scal, r, prec, p, x);
    } else {
      k_gcg_x<T><<<grid_for(n), kBlock, 0, st>>>(n, cg->scal, p, x);   // x current for the host's test
    }
  }
  return check_launch("pyib_grid_cg_iterate");
}

}  // namespace
}  // namespace pyib

extern "C" {
using namespace pyib;

int pyib_grid_flow_mat(const pyib_grid* g, const pyib_work* ws, double rou, double visc, double hdt, const void* p,
                       void* w, double* dot_out, void* stream) {
  auto st = S(stream);
  int nb1 = 0, nb2 = 0;
  if (g->dtype == PYIB_F64)
    gflow<double>(g, ws, rou, visc, hdt, (const double*)p, (double*)w, dot_out != nullptr, &nb1, &nb2, st);
  else if (g->dtype == PYIB_F32)
    gflow<float>(g, ws, rou, visc, hdt, (const float*)p, (float*)w, dot_out != nullptr, &nb1, &nb2, st);
  else
    return fail_msg("pyib_grid_flow_mat: unknown dtype");
  if (dot_out) k_alpha2<<<1, kFinBlock, 0, st>>>(ws->partials, nb1, ws->partials + kRedGrid, nb2, nullptr, 0, dot_out);
  return check_launch("pyib_grid_flow_mat");
}

int pyib_grid_cg_begin(const pyib_grid* g, const pyib_work* ws, const pyib_cg* cg, const void* res, const void* s,
                       double tol1, double tol2, void* stream) {
  auto st = S(stream);
  const int nb = red_grid_for(g->nedge);
  if (g->dtype == PYIB_F64)
    k_gcg_begin<double><<<nb, kBlock, 0, st>>>(g->nedge, (const double*)res, (const double*)s, g->dir,
        (const double*)cg->prec, (double*)cg->r, (double*)cg->x, (double*)cg->p, ws->partials);
  else
    k_gcg_begin<float><<<nb, kBlock, 0, st>>>(g->nedge, (const float*)res, (const float*)s, g->dir,
        (const float*)cg->prec, (float*)cg->r, (float*)cg->x, (float*)cg->p, ws->partials);
  k_gcg_begin_fin<<<1, kFinBlock, 0, st>>>(ws->partials, nb, tol1, tol2, cg->scal);
  return check_launch("pyib_grid_cg_begin");
}

int pyib_grid_cg_iterate(const pyib_grid* g, const pyib_work* ws, const pyib_cg* cg, double rou, double visc,
                         double hdt, int32_t n_iter, int32_t skip_last_direction, void* stream) {
  if (n_iter < 0) return fail_msg("pyib_grid_cg_iterate: negative n_iter");
  if (g->dtype == PYIB_F64)
    return gcg_iterate<double>(g, ws, cg, rou, visc, hdt, n_iter, skip_last_direction, S(stream));
  if (g->dtype == PYIB_F32)
    return gcg_iterate<float>(g, ws, cg, rou, visc, hdt, n_iter, skip_last_direction, S(stream));
  return fail_msg("pyib_grid_cg_iterate: unknown dtype");
}

int pyib_grid_cg_direction(const pyib_grid* g, const pyib_cg* cg, void* stream) {
  auto st = S(stream);
  if (g->dtype == PYIB_F64)
    k_gcg_dir<double><<<grid_for(g->nedge), kBlock, 0, st>>>(g->nedge, cg->scal, (const double*)cg->r,
                                                             (const double*)cg->prec, (double*)cg->p);
  else
    k_gcg_dir<float><<<grid_for(g->nedge), kBlock, 0, st>>>(g->nedge, cg->scal, (const float*)cg->r,
                                                            (const float*)cg->prec, (float*)cg->p);
  k_gcg_shift<<<1, 1, 0, st>>>(cg->scal);
  return check_launch("pyib_grid_cg_direction");
}

int pyib_grid_cg_true_residual(const pyib_grid* g, const pyib_work* ws, const pyib_cg* cg, const void* res,
                               const void* s, void* tmp, void* tr, double rou, double visc, double hdt, double* out2,
                               void* stream) {
  auto st = S(stream);
  const int nb = red_grid_for(g->nedge);
  int nb1, nb2;
  if (g->dtype == PYIB_F64) {
    k_gsub<double><<<grid_for(g->nedge), kBlock, 0, st>>>(g->nedge, (const double*)cg->x, (const double*)s,
                                                          (double*)tmp);
    gflow<double>(g, ws, rou, visc, hdt, (const double*)tmp, (double*)tr, false, &nb1, &nb2, st);
    k_gtrue_res<double><<<nb, kBlock, 0, st>>>(g->nedge, (const double*)res, (const double*)tr, g->dir,
                                               (const double*)cg->prec, (double*)tr, ws->partials);
  } else {
    k_gsub<float><<<grid_for(g->nedge), kBlock, 0, st>>>(g->nedge, (const float*)cg->x, (const float*)s, (float*)tmp);
    gflow<float>(g, ws, rou, visc, hdt, (const float*)tmp, (float*)tr, false, &nb1, &nb2, st);
    k_gtrue_res<float><<<nb, kBlock, 0, st>>>(g->nedge, (const float*)res, (const float*)tr, g->dir,
                                              (const float*)cg->prec, (float*)tr, ws->partials);
  }
  k_finalize<<<1, kFinBlock, 0, st>>>(ws->partials, nb, 2, out2);
  return check_launch("pyib_grid_cg_true_residual");
}

int pyib_cg_replace_residual(double* scal, const double* eta, void* stream) {
  k_gcg_replace<<<1, 1, 0, S(stream)>>>(eta, scal);
  return check_launch("pyib_cg_replace_residual");
}

int pyib_permute(int dtype_in, const void* in, int dtype_out, void* out, int64_t n_out, const int64_t* idx,
                 const int8_t* sigma, void* stream) {
  auto st = S(stream);
  const int gr = grid_for(n_out);
  if (dtype_in == PYIB_F64 && dtype_out == PYIB_F64)
    k_permute<double, double><<<gr, kBlock, 0, st>>>(n_out, idx, sigma, (const double*)in, (double*)out);
  else if (dtype_in == PYIB_F64 && dtype_out == PYIB_F32)
    k_permute<double, float><<<gr, kBlock, 0, st>>>(n_out, idx, sigma, (const double*)in, (float*)out);
  else if (dtype_in == PYIB_F32 && dtype_out == PYIB_F64)
    k_permute<float, double><<<gr, kBlock, 0, st>>>(n_out, idx, sigma, (const float*)in, (double*)out);
  else
    k_permute<float, float><<<gr, kBlock, 0, st>>>(n_out, idx, sigma, (const float*)in, (float*)out);
  return check_launch("pyib_permute");
}

}  // extern "C"
