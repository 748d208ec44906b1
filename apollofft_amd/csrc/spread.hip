// spread.hip -- Kaiser-Bessel spreading (Type-1), interpolation (Type-2) and
// the point-coordinate gradient, 3D subproblems staged in LDS plus 1D paths.
//
// Replaces the per-point T^3 loops of apollo-nufft (type1_into spread,
// transform/dimension_3d.rs:290-327; type2_into gather, :560-601; 1D
// dimension_1d.rs:186-197, :261-275) and the O(grid x points) WGSL passes
// fast_type1_spread_3d / fast_type2_interpolate_3d (apollo-nufft-wgpu
// .../shaders/nufft_fast_3d.wgsl:130-161, :268-299).
//
// 3D work unit = subproblem: up to max_sub points of one tile.
//
// Spreading (Type-1) is output-stationary: a 1024-thread workgroup owns the
// padded tile of a subproblem, lane (py, pz) of the workgroup owns one (y,z)
// column of it (32 x 32 columns; each wave owns an 8 x 8 block of columns)
// and keeps the column's bin_x + 2w x-cells in registers.  Points are staged
// 256 at a time in LDS (coordinates -> grid units in FP64, 2w FP32 tap
// weights per axis from the per-tap polynomials, value, tap origin); every
// wave walks the staged points and skips, with one uniform test, those whose
// (y,z) support misses its column block; otherwise each lane multiplies the
// point's value by its (y,z) weight and adds the 2w x-taps into its
// registers (the x origin is uniform, so a switch selects statically indexed
// registers).  Nothing is accumulated through LDS atomics: on gfx950 the
// LDS FP32 atomic add measured 0.33 lane-ops/clk/CU, FP64 7.8, against ~64
// lane-FMAs/clk/CU for register FMAs.  At the end each lane adds its column
// to the HBM grid with global float atomics (periodic wrap, zeros skipped).
//
// Interpolation (Type-2) loads its own, smaller padded tile from HBM into
// LDS once per subproblem; each wave gathers one point at a time (lanes over
// (y,z) taps, loop over x) and reduces across the wave.
#include "internal.h"

namespace anufft {
namespace {

constexpr int kStage = 128;        // points staged per round (interp, grad)
constexpr int kSpreadStage = 256;  // points staged per round (spread)
constexpr int kSpreadThreads = 1024;
constexpr int kInterpThreads = 1024;

// Periodic index; padded tiles may exceed the grid by more than one period
// when the grid is not much larger than the kernel support.
__device__ __forceinline__ int wrap(int v, int m) {
  while (v < 0) v += m;
  while (v >= m) v -= m;
  return v;
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}

// Sum over the 64 lanes of a wave with DPP (quad swaps, row mirrors, row
// broadcasts) instead of LDS permutes; the total is read from lane 63.
__device__ __forceinline__ float wave_sum(float v) {
  v += dpp<0xB1>(v);        // quad_perm [1,0,3,2]
  v += dpp<0x4E>(v);        // quad_perm [2,3,0,1]
  v += dpp<0x141>(v);       // row_half_mirror
  v += dpp<0x140>(v);       // row_mirror: every lane holds its row's sum
  v += dpp<0x142, 0xa>(v);  // row_bcast15 into rows 1 and 3
  v += dpp<0x143, 0xc>(v);  // row_bcast31 into rows 2 and 3
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

template <int W>
struct Staging {
  static constexpr int T = 2 * W;
  static constexpr int WS = (3 * T + 3) / 4 * 4; // weights per point, 16-byte rows
};

// Stage one round of points: weights, values, tap origins (3 tasks per point).
template <int W, int NT, bool DERIV>
__device__ __forceinline__ void stage_points(const PointArgs &pa, const Geom &g,
                                             const KbPoly &poly, const float2 *values,
                                             int start, int npts, int org0, int org1, int org2,
                                             float *wts, float *dwts, float2 *cv, int *loc,
                                             int *jidx) {
  constexpr int T = 2 * W;
  constexpr int WS = Staging<W>::WS;
  for (int task = threadIdx.x; task < 3 * npts; task += NT) {
    int ax = task / npts;
    int p = task - ax * npts;
    uint32_t j = pa.perm[start + p];
    const float *cp = ax == 0 ? pa.x : (ax == 1 ? pa.y : pa.z);
    int cell;
    double t = fold_coord(cp[(int64_t)j * pa.stride], g.scale[ax], g.m[ax], cell);
    float wt[T];
    int i0;
    if (DERIV) {
      float dwt[T];
      i0 = kb_taps_deriv<W>(t, cell, poly, wt, dwt);
#pragma unroll
      for (int i = 0; i < T; ++i) dwts[p * WS + ax * T + i] = dwt[i];
    } else {
      i0 = kb_taps<W>(t, cell, poly, wt);
    }
#pragma unroll
    for (int i = 0; i < T; ++i) wts[p * WS + ax * T + i] = wt[i];
    int org = ax == 0 ? org0 : (ax == 1 ? org1 : org2);
    loc[p * 3 + ax] = i0 - org;
    if (ax == 0) {
      if (values) cv[p] = values[j];
      if (jidx) jidx[p] = (int)j;
    }
  }
}

// Per-point spreading record, in sorted order: the value (re, im) and the 2w
// x-weights, zero-padded to a multiple of 4 floats.  The spreading kernel
// reads a record with wave-uniform scalar loads, so the x-weights reach the
// FMAs as scalar operands instead of 64-lane LDS broadcasts.
template <int W>
struct XRec {
  static constexpr int T = 2 * W;
  static constexpr int RW = (T + 2 + 3) / 4 * 4;  // floats per record
};

// Fill the records and the x offset cell - bin_x floor(cell / bin_x) of the
// first tap's register cell, for every sorted position.
template <int W>
__global__ __launch_bounds__(256) void spread_xrec_kernel(PointArgs pa, const float2 *values,
                                                           Geom g, KbPoly poly, float4 *rec,
                                                           uint8_t *xo) {
  constexpr int T = 2 * W;
  constexpr int RW = XRec<W>::RW;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < pa.num_points; s += st) {
    const uint32_t j = pa.perm[s];
    int cell;
    const double t = fold_coord(pa.x[(int64_t)j * pa.stride], g.scale[0], g.m[0], cell);
    float r[RW];
    kb_taps<W>(t, cell, poly, r + 2);
    const float2 c = values[j];
    r[0] = c.x;
    r[1] = c.y;
#pragma unroll
    for (int i = T + 2; i < RW; ++i) r[i] = 0.f;
    float4 *d = rec + s * (RW / 4);
#pragma unroll
    for (int q = 0; q < RW / 4; ++q) d[q] = make_float4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
    xo[s] = (uint8_t)(cell - (cell / g.bin[0]) * g.bin[0]);
  }
}

// LDS staging of the y and z weights of one point, zero-padded to the 32
// column positions of the padded tile, so a lane reads its weight at its own
// column index.
constexpr int kSpreadRow = 2 * kSpreadCols;  // floats per staged point

// Stage points for spreading: 2 tasks per point (y and z).  Each zero-fills
// its row, stores the 2w weights at the taps' column, and writes a 4-bit
// mask of the 8-column lane blocks the point's support touches; the y task
// also copies the point's x offset.
template <int W, int NT>
__device__ __forceinline__ void stage_spread(const PointArgs &pa, const Geom &g,
                                             const float *pc, const uint8_t *xo, int start,
                                             int npts, int oy, int oz, float *rows, int *rowm,
                                             int *colm, int *xos) {
  constexpr int T = 2 * W;
  for (int task = threadIdx.x; task < 2 * npts; task += NT) {
    const int ax = task < npts ? 1 : 2;
    const int p = task - (ax - 1) * npts;
    const uint32_t j = pa.perm[start + p];
    const float *cp = ax == 1 ? pa.y : pa.z;
    int cell;
    const double t = fold_coord(cp[(int64_t)j * pa.stride], g.scale[ax], g.m[ax], cell);
    const int l = cell - W + 1 - (ax == 1 ? oy : oz);  // first tap's column, in [1, bin]
    float *r = rows + p * kSpreadRow + (ax - 1) * kSpreadCols;
#pragma unroll
    for (int i = 0; i < kSpreadCols; i += 4) *(float4 *)(r + i) = make_float4(0.f, 0.f, 0.f, 0.f);
    // the taps of kb_taps, stored one by one, with the polynomial
    // coefficients read from the workgroup's LDS copy pc[d][i]
    const float z = (float)(1.0 - 2.0 * (t - (double)cell));
#pragma unroll 1
    for (int i = 0; i < T; ++i) {
      float v = pc[(kPolyCoefs - 1) * kMaxTaps + i];
#pragma unroll
      for (int d = kPolyCoefs - 2; d >= 0; --d) v = fmaf(v, z, pc[d * kMaxTaps + i]);
      if (i == 0 && z <= -1.0f) v = 0.0f;
      if (i == T - 1 && z >= 1.0f) v = 0.0f;
      r[l + i] = v;
    }
    const int m = ((1 << ((l + T - 1) / 8 + 1)) - 1) & ~((1 << (l / 8)) - 1);
    if (ax == 1) {
      rowm[p] = m;
      xos[p] = xo[start + p];
    } else {
      colm[p] = m;
    }
  }
}

typedef const float __attribute__((address_space(4))) *ConstFloatPtr;

// One spreading subproblem (see the file comment): stage, accumulate the
// lane's column in registers, flush it to the grid.
// LDS of the spreading kernels: staged rows, masks and x offsets, then the
// polynomial coefficients (copied once per workgroup by spread_load_poly).
constexpr int kSpreadPolyOfs = kSpreadStage * (kSpreadRow + 3);  // floats

__device__ __forceinline__ void spread_load_poly(const KbPoly &poly, float *sh) {
  float *pc = sh + kSpreadPolyOfs;
  for (int i = threadIdx.x; i < kPolyCoefs * kMaxTaps; i += blockDim.x)
    pc[i] = poly.c[i / kMaxTaps][i % kMaxTaps];
}

template <int W>
__device__ __forceinline__ void spread_subproblem(const PointArgs &pa, const float *rec,
                                                  const uint8_t *xo, float2 *grid, const Geom &g,
                                                  int4 sub, float *sh) {
  constexpr int NT = kSpreadThreads;
  constexpr int T = 2 * W;
  constexpr int PX = kSpreadBinX + T - 1;  // x cells of a tile column
  constexpr int RW = XRec<W>::RW;
  float *rows = sh;
  int *rowm = (int *)(rows + kSpreadStage * kSpreadRow);
  int *colm = rowm + kSpreadStage;
  int *xos = colm + kSpreadStage;
  const float *pc = sh + kSpreadPolyOfs;

  // Wave -> 8 x 8 column block.  Waves are dealt to the 4 SIMDs round-robin;
  // the z block is rotated by the y block so each SIMD holds one block of
  // every row and column of blocks (central blocks see the most points).
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wyb = wave >> 2, wzb = (wave + wyb) & 3;
  const int py = wyb * 8 + (lane >> 3), pz = wzb * 8 + (lane & 7);
  const bool col_ok = py < g.pad[1] && pz < g.pad[2];

  const int bin = sub.x;
  const int bz = bin % g.nbin[2];
  const int by = (bin / g.nbin[2]) % g.nbin[1];
  const int bx = bin / (g.nbin[2] * g.nbin[1]);
  const int bx0 = bx * g.bin[0];
  const int oy = by * g.bin[1] - W, oz = bz * g.bin[2] - W;

  float ar[PX], ai[PX];
#pragma unroll
  for (int i = 0; i < PX; ++i) {
    ar[i] = 0.f;
    ai[i] = 0.f;
  }
  for (int c0 = 0; c0 < sub.z; c0 += kSpreadStage) {
    const int npts = min(kSpreadStage, sub.z - c0);
    __syncthreads();
    stage_spread<W, NT>(pa, g, pc, xo, sub.y + c0, npts, oy, oz, rows, rowm, colm, xos);
    __syncthreads();
    const ConstFloatPtr recs = (ConstFloatPtr)(rec + (int64_t)(sub.y + c0) * RW);
    for (int p0 = 0; p0 < npts; p0 += 64) {
      const int q = p0 + lane;
      const bool hit = q < npts && ((rowm[q] >> wyb) & (colm[q] >> wzb) & 1);
      const int o = q < npts ? xos[q] : 0;
      // one pass per x offset, so the register cells it adds to are static
#pragma unroll
      for (int oo = 0; oo < kSpreadBinX; ++oo) {
        uint64_t mask = __ballot(hit && o == oo);
        while (mask) {
          const int k = __ffsll((unsigned long long)mask) - 1;
          mask &= mask - 1;
          const int p = p0 + k;
          const float *r = rows + p * kSpreadRow;
          const float wyz = r[py] * r[kSpreadCols + pz];
          const ConstFloatPtr xr = recs + p * RW;
          const float cr = xr[0] * wyz, ci = xr[1] * wyz;
#pragma unroll
          for (int i = 0; i < T; ++i) {
            const float wx = xr[2 + i];
            ar[oo + i] = fmaf(cr, wx, ar[oo + i]);
            ai[oo + i] = fmaf(ci, wx, ai[oo + i]);
          }
        }
      }
    }
  }

  // add this lane's column to the grid: cell i <-> x = bx0 - W + 1 + i
  if (col_ok) {
    const int gy = wrap(oy + py, g.m[1]);
    const int gz = wrap(oz + pz, g.m[2]);
#pragma unroll
    for (int i = 0; i < PX; ++i) {
      if (ar[i] != 0.f || ai[i] != 0.f) {
        const int gx = wrap(bx0 - W + 1 + i, g.m[0]);
        float2 *dst = grid + ((int64_t)gx * g.m[1] + gy) * g.m[2] + gz;
        unsafeAtomicAdd(&dst->x, ar[i]);
        unsafeAtomicAdd(&dst->y, ai[i]);
      }
    }
  }
}

template <int W>
__global__ __launch_bounds__(kSpreadThreads) void spread3d_kernel(PointArgs pa, const float *rec,
                                                                  const uint8_t *xo, float2 *grid,
                                                                  Geom g, KbPoly poly) {
  extern __shared__ float sh[];
  spread_load_poly(poly, sh);  // ordered before use by the staging barrier
  const uint32_t nsub = *pa.nsub;
  for (uint32_t sidx = blockIdx.x; sidx < nsub; sidx += gridDim.x)
    spread_subproblem<W>(pa, rec, xo, grid, g, pa.subs[sidx], sh);
}

// Batched spreading: one persistent launch over the subproblems of all
// problems.  prefix[p] is the first global subproblem index of problem p
// (prefix[nprob] = total); the problem of a global index is found by a
// uniform binary search and its descriptor read with scalar loads.
template <int W>
__global__ __launch_bounds__(kSpreadThreads) void spread3d_batched_kernel(
    const SpreadProblem *probs, const uint32_t *prefix, int nprob, KbPoly poly) {
  extern __shared__ float sh[];
  spread_load_poly(poly, sh);  // ordered before use by the staging barrier
  const uint32_t total = prefix[nprob];
  for (uint32_t gidx = blockIdx.x; gidx < total; gidx += gridDim.x) {
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (prefix[mid] <= gidx) lo = mid;
      else hi = mid - 1;
    }
    const SpreadProblem &pr = probs[lo];
    spread_subproblem<W>(pr.pa, pr.rec, pr.xo, pr.grid, pr.g, pr.pa.subs[gidx - prefix[lo]], sh);
  }
}

// Write descriptors passed by value into the device table.
__global__ void batch_table_kernel(BatchChunk chunk, SpreadProblem *table) {
  const int i = threadIdx.x;
  if (i < chunk.count) table[chunk.first + i] = chunk.probs[i];
}

// prefix[p] = sum of the subproblem counts of problems < p.
__global__ void batch_prefix_kernel(const SpreadProblem *probs, int nprob, uint32_t *prefix) {
  if (threadIdx.x != 0) return;
  uint32_t acc = 0;
  for (int p = 0; p < nprob; ++p) {
    prefix[p] = acc;
    acc += *probs[p].pa.nsub;
  }
  prefix[nprob] = acc;
}

// Load the padded tile of subproblem origin (ox,oy,oz) from the grid.
__device__ __forceinline__ void load_tile(const float2 *grid, const Geom &g, int ox, int oy,
                                          int oz, float2 *tile, int NT) {
  const int Px = g.pad[0], Py = g.pad[1], Pz = g.pad[2];
  const int P = Px * Py * Pz;
  for (int i = threadIdx.x; i < P; i += NT) {
    const int pz = i % Pz;
    const int t2 = i / Pz;
    const int py = t2 % Py;
    const int px = t2 / Py;
    const int gx = wrap(ox + px, g.m[0]);
    const int gy = wrap(oy + py, g.m[1]);
    const int gz = wrap(oz + pz, g.m[2]);
    tile[i] = grid[((int64_t)gx * g.m[1] + gy) * g.m[2] + gz];
  }
}

template <int W>
__global__ __launch_bounds__(kInterpThreads) void interp3d_kernel(PointArgs pa,
                                                                  const float2 *grid,
                                                                  float2 *out, Geom g,
                                                                  KbPoly poly) {
  constexpr int T = 2 * W;
  constexpr int WS = Staging<W>::WS;
  constexpr int NT = kInterpThreads;
  constexpr int kWaves = NT / 64;
  constexpr int kRounds = (T * T + 63) / 64;
  extern __shared__ float sh[];
  const int Py = g.pad[1], Pz = g.pad[2];
  const int P = g.pad[0] * Py * Pz;
  const int PyPz = Py * Pz;
  float2 *tile = (float2 *)sh;
  float *wts = (float *)(tile + P);
  int *loc = (int *)(wts + kStage * WS);
  int *jidx = loc + 3 * kStage;

  const uint32_t nsub = *pa.nsub;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t sidx = blockIdx.x; sidx < nsub; sidx += gridDim.x) {
    const int4 sub = pa.subs[sidx];
    const int bin = sub.x;
    const int bz = bin % g.nbin[2];
    const int by = (bin / g.nbin[2]) % g.nbin[1];
    const int bx = bin / (g.nbin[2] * g.nbin[1]);
    const int ox = bx * g.bin[0] - W, oy = by * g.bin[1] - W, oz = bz * g.bin[2] - W;
    __syncthreads();
    load_tile(grid, g, ox, oy, oz, tile, NT);
    for (int c0 = 0; c0 < sub.z; c0 += kStage) {
      const int npts = min(kStage, sub.z - c0);
      __syncthreads();
      stage_points<W, NT, false>(pa, g, poly, nullptr, sub.y + c0, npts, ox, oy, oz, wts,
                                 nullptr, nullptr, loc, jidx);
      __syncthreads();
      for (int p = wave; p < npts; p += kWaves) {
        const float *wp = wts + p * WS;
        const int lx = loc[3 * p], ly = loc[3 * p + 1], lz = loc[3 * p + 2];
        float ar = 0.f, ai = 0.f;
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
          const int q = lane + 64 * r;
          if (q < T * T) {
            const int ty = q / T, tz = q - (q / T) * T;
            const float wyz = wp[T + ty] * wp[2 * T + tz];
            const float2 *tp = tile + (lx * Py + ly + ty) * Pz + lz + tz;
            float sr = 0.f, si = 0.f;
#pragma unroll
            for (int tx = 0; tx < T; ++tx) {
              const float2 v = tp[tx * PyPz];
              sr = fmaf(wp[tx], v.x, sr);
              si = fmaf(wp[tx], v.y, si);
            }
            ar = fmaf(wyz, sr, ar);
            ai = fmaf(wyz, si, ai);
          }
        }
        ar = wave_sum(ar);
        ai = wave_sum(ai);
        if (lane == 0) out[jidx[p]] = make_float2(ar, ai);
      }
    }
  }
}

// d/dx of the interpolant, contracted with conj(u_j):
// grad[a][j] = Re(conj(u_j) * (-scale_a) * sum dphi_a phi_b phi_c G).
template <int W>
__global__ __launch_bounds__(kInterpThreads) void grad3d_kernel(PointArgs pa,
                                                                const float2 *grid,
                                                                const float2 *u, float *gx,
                                                                float *gy, float *gz, Geom g,
                                                                KbPoly poly) {
  constexpr int T = 2 * W;
  constexpr int WS = Staging<W>::WS;
  constexpr int NT = kInterpThreads;
  constexpr int kWaves = NT / 64;
  constexpr int kRounds = (T * T + 63) / 64;
  extern __shared__ float sh[];
  const int Py = g.pad[1], Pz = g.pad[2];
  const int P = g.pad[0] * Py * Pz;
  const int PyPz = Py * Pz;
  float2 *tile = (float2 *)sh;
  float *wts = (float *)(tile + P);
  float *dwts = wts + kStage * WS;
  int *loc = (int *)(dwts + kStage * WS);
  int *jidx = loc + 3 * kStage;

  const uint32_t nsub = *pa.nsub;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t sidx = blockIdx.x; sidx < nsub; sidx += gridDim.x) {
    const int4 sub = pa.subs[sidx];
    const int bin = sub.x;
    const int bz = bin % g.nbin[2];
    const int by = (bin / g.nbin[2]) % g.nbin[1];
    const int bx = bin / (g.nbin[2] * g.nbin[1]);
    const int ox = bx * g.bin[0] - W, oy = by * g.bin[1] - W, oz = bz * g.bin[2] - W;
    __syncthreads();
    load_tile(grid, g, ox, oy, oz, tile, NT);
    for (int c0 = 0; c0 < sub.z; c0 += kStage) {
      const int npts = min(kStage, sub.z - c0);
      __syncthreads();
      stage_points<W, NT, true>(pa, g, poly, nullptr, sub.y + c0, npts, ox, oy, oz, wts, dwts,
                                nullptr, loc, jidx);
      __syncthreads();
      for (int p = wave; p < npts; p += kWaves) {
        const float *wp = wts + p * WS;
        const float *dp = dwts + p * WS;
        const int lx = loc[3 * p], ly = loc[3 * p + 1], lz = loc[3 * p + 2];
        float sx[2] = {0.f, 0.f}, sy[2] = {0.f, 0.f}, sz[2] = {0.f, 0.f};
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
          const int q = lane + 64 * r;
          if (q < T * T) {
            const int ty = q / T, tz = q - (q / T) * T;
            const float2 *tp = tile + (lx * Py + ly + ty) * Pz + lz + tz;
            float ar = 0.f, ai = 0.f, dr = 0.f, di = 0.f;
#pragma unroll
            for (int tx = 0; tx < T; ++tx) {
              const float2 v = tp[tx * PyPz];
              ar = fmaf(wp[tx], v.x, ar);
              ai = fmaf(wp[tx], v.y, ai);
              dr = fmaf(dp[tx], v.x, dr);
              di = fmaf(dp[tx], v.y, di);
            }
            const float wy = wp[T + ty], wz = wp[2 * T + tz];
            const float dy = dp[T + ty], dz = dp[2 * T + tz];
            const float wyz = wy * wz, dyz = dy * wz, wydz = wy * dz;
            sx[0] = fmaf(wyz, dr, sx[0]);
            sx[1] = fmaf(wyz, di, sx[1]);
            sy[0] = fmaf(dyz, ar, sy[0]);
            sy[1] = fmaf(dyz, ai, sy[1]);
            sz[0] = fmaf(wydz, ar, sz[0]);
            sz[1] = fmaf(wydz, ai, sz[1]);
          }
        }
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          sx[c] = wave_sum(sx[c]);
          sy[c] = wave_sum(sy[c]);
          sz[c] = wave_sum(sz[c]);
        }
        if (lane == 0) {
          const int j = jidx[p];
          const float2 uj = u[j];
          // Re(conj(u) * (-s) * S) = -s (u.x S.x + u.y S.y)
          gx[j] = -(float)g.scale[0] * (uj.x * sx[0] + uj.y * sx[1]);
          if (gy) gy[j] = -(float)g.scale[1] * (uj.x * sy[0] + uj.y * sy[1]);
          if (gz) gz[j] = -(float)g.scale[2] * (uj.x * sz[0] + uj.y * sz[1]);
        }
      }
    }
  }
}

// ---- 1D: one thread per point, global atomics / gathers ---------------------
template <int W>
__global__ void spread1d_kernel(const float *x, int64_t stride, int64_t n,
                                const float2 *values, float2 *grid, Geom g, KbPoly poly) {
  constexpr int T = 2 * W;
  const int m = g.m[0];
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x) {
    int cell;
    double t = fold_coord(x[j * stride], g.scale[0], m, cell);
    float wt[T];
    int i0 = kb_taps<W>(t, cell, poly, wt);
    const float2 c = values[j];
#pragma unroll
    for (int i = 0; i < T; ++i) {
      int idx = (i0 + i) % m;
      idx = idx < 0 ? idx + m : idx;
      unsafeAtomicAdd(&grid[idx].x, c.x * wt[i]);
      unsafeAtomicAdd(&grid[idx].y, c.y * wt[i]);
    }
  }
}

template <int W>
__global__ void interp1d_kernel(const float *x, int64_t stride, int64_t n, const float2 *grid,
                                float2 *out, Geom g, KbPoly poly) {
  constexpr int T = 2 * W;
  const int m = g.m[0];
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x) {
    int cell;
    double t = fold_coord(x[j * stride], g.scale[0], m, cell);
    float wt[T];
    int i0 = kb_taps<W>(t, cell, poly, wt);
    float ar = 0.f, ai = 0.f;
#pragma unroll
    for (int i = 0; i < T; ++i) {
      int idx = (i0 + i) % m;
      idx = idx < 0 ? idx + m : idx;
      const float2 v = grid[idx];
      ar = fmaf(wt[i], v.x, ar);
      ai = fmaf(wt[i], v.y, ai);
    }
    out[j] = make_float2(ar, ai);
  }
}

template <int W>
__global__ void grad1d_kernel(const float *x, int64_t stride, int64_t n, const float2 *grid,
                              const float2 *u, float *gx, Geom g, KbPoly poly) {
  constexpr int T = 2 * W;
  const int m = g.m[0];
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x) {
    int cell;
    double t = fold_coord(x[j * stride], g.scale[0], m, cell);
    float wt[T], dwt[T];
    int i0 = kb_taps_deriv<W>(t, cell, poly, wt, dwt);
    float dr = 0.f, di = 0.f;
#pragma unroll
    for (int i = 0; i < T; ++i) {
      int idx = (i0 + i) % m;
      idx = idx < 0 ? idx + m : idx;
      const float2 v = grid[idx];
      dr = fmaf(dwt[i], v.x, dr);
      di = fmaf(dwt[i], v.y, di);
    }
    const float2 uj = u[j];
    gx[j] = -(float)g.scale[0] * (uj.x * dr + uj.y * di);
  }
}

template <int W>
size_t spread_lds(const Geom &) {
  return ((size_t)kSpreadPolyOfs + kPolyCoefs * kMaxTaps) * sizeof(float);
}

template <int W>
size_t interp_lds(const Geom &g, bool grad) {
  const size_t P = (size_t)g.pad[0] * g.pad[1] * g.pad[2];
  const size_t ws = (size_t)kStage * Staging<W>::WS * (grad ? 2 : 1);
  return P * sizeof(float2) + ws * sizeof(float) + 4 * kStage * sizeof(int);
}

template <template <int> class F, typename... A>
hipError_t dispatch_w(int w, A &&...args) {
  switch (w) {
  case 2: return F<2>::run(args...);
  case 3: return F<3>::run(args...);
  case 4: return F<4>::run(args...);
  case 5: return F<5>::run(args...);
  case 6: return F<6>::run(args...);
  case 7: return F<7>::run(args...);
  case 8: return F<8>::run(args...);
  default: return hipErrorInvalidValue;
  }
}

template <int W>
struct SpreadXrec {
  static hipError_t run(const PointArgs &pa, const float2 *values, const Geom &g,
                        const KbPoly &poly, float *rec, uint8_t *xo, hipStream_t s) {
    if (pa.num_points <= 0) return hipSuccess;
    int64_t blocks = (pa.num_points + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(spread_xrec_kernel<W>, dim3((unsigned)blocks), dim3(256), 0, s, pa, values,
                       g, poly, (float4 *)rec, xo);
    return hipGetLastError();
  }
};

template <int W>
struct Spread3 {
  static hipError_t run(const PointArgs &pa, const float *rec, const uint8_t *xo, float2 *grid,
                        const Geom &g, const KbPoly &poly, int nblk, hipStream_t s) {
    size_t lds = spread_lds<W>(g);
    hipError_t e = hipFuncSetAttribute((const void *)spread3d_kernel<W>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(spread3d_kernel<W>, dim3(nblk), dim3(kSpreadThreads), lds, s, pa, rec, xo,
                       grid, g, poly);
    return hipGetLastError();
  }
};

template <int W>
struct Spread3Batched {
  static hipError_t run(const SpreadProblem *probs, const uint32_t *prefix, int nprob,
                        const KbPoly &poly, int nblk, hipStream_t s) {
    const Geom g{};
    size_t lds = spread_lds<W>(g);
    hipError_t e = hipFuncSetAttribute((const void *)spread3d_batched_kernel<W>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(spread3d_batched_kernel<W>, dim3(nblk), dim3(kSpreadThreads), lds, s,
                       probs, prefix, nprob, poly);
    return hipGetLastError();
  }
};

template <int W>
struct Interp3 {
  static hipError_t run(const PointArgs &pa, const float2 *grid, float2 *out, const Geom &g,
                        const KbPoly &poly, int nblk, hipStream_t s) {
    size_t lds = interp_lds<W>(g, false);
    hipError_t e = hipFuncSetAttribute((const void *)interp3d_kernel<W>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(interp3d_kernel<W>, dim3(nblk), dim3(kInterpThreads), lds, s, pa, grid,
                       out, g, poly);
    return hipGetLastError();
  }
};

template <int W>
struct Grad3 {
  static hipError_t run(const PointArgs &pa, const float2 *grid, const float2 *u, float *gx,
                        float *gy, float *gz, const Geom &g, const KbPoly &poly, int nblk,
                        hipStream_t s) {
    size_t lds = interp_lds<W>(g, true);
    hipError_t e = hipFuncSetAttribute((const void *)grad3d_kernel<W>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(grad3d_kernel<W>, dim3(nblk), dim3(kInterpThreads), lds, s, pa, grid, u,
                       gx, gy, gz, g, poly);
    return hipGetLastError();
  }
};

inline unsigned blocks_1d(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > 65536) b = 65536;
  return (unsigned)b;
}

template <int W>
struct Spread1 {
  static hipError_t run(const float *x, int64_t stride, int64_t n, const float2 *values,
                        float2 *grid, const Geom &g, const KbPoly &poly, hipStream_t s) {
    hipLaunchKernelGGL(spread1d_kernel<W>, dim3(blocks_1d(n)), dim3(256), 0, s, x, stride, n,
                       values, grid, g, poly);
    return hipGetLastError();
  }
};

template <int W>
struct Interp1 {
  static hipError_t run(const float *x, int64_t stride, int64_t n, const float2 *grid,
                        float2 *out, const Geom &g, const KbPoly &poly, hipStream_t s) {
    hipLaunchKernelGGL(interp1d_kernel<W>, dim3(blocks_1d(n)), dim3(256), 0, s, x, stride, n,
                       grid, out, g, poly);
    return hipGetLastError();
  }
};

template <int W>
struct Grad1 {
  static hipError_t run(const float *x, int64_t stride, int64_t n, const float2 *grid,
                        const float2 *u, float *gx, const Geom &g, const KbPoly &poly,
                        hipStream_t s) {
    hipLaunchKernelGGL(grad1d_kernel<W>, dim3(blocks_1d(n)), dim3(256), 0, s, x, stride, n,
                       grid, u, gx, g, poly);
    return hipGetLastError();
  }
};

template <int W>
struct LdsSize {
  static hipError_t run(const Geom &g, int kind, size_t *out) {
    *out = kind == 0 ? spread_lds<W>(g) : interp_lds<W>(g, kind == 2);
    return hipSuccess;
  }
};

} // namespace

int spread_rec_floats(int w) { return (2 * w + 2 + 3) / 4 * 4; }

hipError_t launch_spread_xrec(int w, const PointArgs &pa, const float2 *values, const Geom &g,
                              const KbPoly &poly, float *rec, uint8_t *xo, hipStream_t s) {
  return dispatch_w<SpreadXrec>(w, pa, values, g, poly, rec, xo, s);
}

hipError_t launch_spread3d(int w, const PointArgs &pa, const float *rec, const uint8_t *xo,
                           float2 *grid, const Geom &g, const KbPoly &poly, int nblk,
                           hipStream_t s) {
  return dispatch_w<Spread3>(w, pa, rec, xo, grid, g, poly, nblk, s);
}
hipError_t launch_spread3d_batched(int w, const SpreadProblem *host_probs, int nprob,
                                   SpreadProblem *dev_table, uint32_t *dev_prefix,
                                   const KbPoly &poly, int nblk, hipStream_t s) {
  for (int first = 0; first < nprob; first += kBatchChunk) {
    BatchChunk chunk{};
    chunk.first = first;
    chunk.count = nprob - first < kBatchChunk ? nprob - first : kBatchChunk;
    for (int i = 0; i < chunk.count; ++i) chunk.probs[i] = host_probs[first + i];
    hipLaunchKernelGGL(batch_table_kernel, dim3(1), dim3(64), 0, s, chunk, dev_table);
  }
  hipLaunchKernelGGL(batch_prefix_kernel, dim3(1), dim3(64), 0, s, dev_table, nprob, dev_prefix);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return dispatch_w<Spread3Batched>(w, (const SpreadProblem *)dev_table, (const uint32_t *)dev_prefix,
                                    nprob, poly, nblk, s);
}

hipError_t launch_interp3d(int w, const PointArgs &pa, const float2 *grid, float2 *out,
                           const Geom &g, const KbPoly &poly, int nblk, hipStream_t s) {
  return dispatch_w<Interp3>(w, pa, grid, out, g, poly, nblk, s);
}
hipError_t launch_grad3d(int w, const PointArgs &pa, const float2 *grid, const float2 *u,
                         float *gx, float *gy, float *gz, const Geom &g, const KbPoly &poly,
                         int nblk, hipStream_t s) {
  return dispatch_w<Grad3>(w, pa, grid, u, gx, gy, gz, g, poly, nblk, s);
}
hipError_t launch_spread1d(int w, const float *x, int64_t stride, int64_t n,
                           const float2 *values, float2 *grid, const Geom &g,
                           const KbPoly &poly, hipStream_t s) {
  return dispatch_w<Spread1>(w, x, stride, n, values, grid, g, poly, s);
}
hipError_t launch_interp1d(int w, const float *x, int64_t stride, int64_t n,
                           const float2 *grid, float2 *out, const Geom &g, const KbPoly &poly,
                           hipStream_t s) {
  return dispatch_w<Interp1>(w, x, stride, n, grid, out, g, poly, s);
}
hipError_t launch_grad1d(int w, const float *x, int64_t stride, int64_t n, const float2 *grid,
                         const float2 *u, float *gx, const Geom &g, const KbPoly &poly,
                         hipStream_t s) {
  return dispatch_w<Grad1>(w, x, stride, n, grid, u, gx, g, poly, s);
}
size_t spread_lds_bytes(int w, const Geom &g) {
  size_t v = 0;
  (void)dispatch_w<LdsSize>(w, g, 0, &v);
  return v;
}
size_t interp_lds_bytes(int w, const Geom &g, bool grad) {
  size_t v = 0;
  (void)dispatch_w<LdsSize>(w, g, grad ? 2 : 1, &v);
  return v;
}

} // namespace anufft
