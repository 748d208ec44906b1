// spread.hip -- Kaiser-Bessel spreading (Type-1), interpolation (Type-2) and
// the point-coordinate gradient: 3D subproblems (up to max_sub points of one
// tile) on the matrix cores or the vector ALUs, plus the 1D paths.
//
// Replaces the per-point T^3 loops of apollo-nufft (type1_into spread,
// transform/dimension_3d.rs:290-327; type2_into gather, :560-601; 1D
// dimension_1d.rs:186-197, :261-275) and the O(grid x points) WGSL passes
// fast_type1_spread_3d / fast_type2_interpolate_3d (apollo-nufft-wgpu
// .../shaders/nufft_fast_3d.wgsl:130-161, :268-299).
//
// Matrix-core kernels (default): a tile is PX x cells x 32 y x 32 z.  For
// each x cell the (y,z) plane is one 32 x 32 output of
// v_mfma_f32_32x32x16_f16 over 16 points: spreading multiplies
// A[y][k] = wx_k wy_k(y) by B[k][z] = c_k wz_k(z); interpolation multiplies
// the grid plane by wz_k(z) and contracts y and x on the vector ALUs.  FP32
// operands are split into FP16 hi + lo (three products per FP32 product,
// FP32 accumulation), with power-of-two scales per subproblem / tile.
// Points come from 32-byte records written in sorted order by the binning.
// A workgroup walks a contiguous, work-balanced run of the subproblem list
// (pencil order); spreading keeps its accumulators as a ring of x columns
// and flushes only the columns a tile leaves behind.
//
// Vector-ALU kernels: spreading is output-stationary: a 1024-thread
// workgroup owns the padded tile of a subproblem, lane (py, pz) owns one
// (y,z) column and keeps the column's bin_x + 2w x-cells in registers;
// points are staged 256 at a time in LDS and each lane adds the 2w x-taps of
// the points whose support meets its column.  Nothing is accumulated through
// LDS atomics: on gfx950 the LDS FP32 atomic add measured 0.33
// lane-ops/clk/CU, FP64 7.8, against ~64 lane-FMAs/clk/CU for register
// FMAs.  Interpolation loads its padded tile into LDS once per subproblem;
// each wave gathers one point at a time and reduces across the wave.
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
    double t = fold_coord(cp[(int64_t)j * pa.stride], g.scale[ax], g.m[ax], g.inv_m[ax], cell);
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

// Per-point record, in sorted order (floats):
//   [0, 2)          the value (re, im); zero for Type-2
//   [2, 2 + 2w)     x tap weights,  [2 + 2w, 2 + 4w) y,  [2 + 4w, 2 + 6w) z
//   2 + 6w          tile offsets of the first taps (bit pattern of an int:
//                   x register cell | y column << 8 | z column << 16)
// zero-padded to a multiple of 4.  Computing every weight once, in one fully
// parallel pass, leaves the tile kernels only coalesced copies to stage.  The
// vector-ALU spreading kernel reads the value and x weights of a record with
// wave-uniform scalar loads, so they reach its FMAs as scalar operands.
template <int W>
struct XRec {
  static constexpr int T = 2 * W;
  static constexpr int WX = 2, WY = 2 + T, WZ = 2 + 2 * T, OFF = 2 + 3 * T;
  static constexpr int RW = (OFF + 1 + 3) / 4 * 4;  // floats per record
};

// Fill the record and the x offset (cell mod bin_x) of every sorted point.
// Offsets: x register cell of the first tap = cell_x mod bin_x; y (z) tile
// column of the first tap = cell mod bin + 1 (the tile starts w cells before
// its bin).
template <int W>
__global__ __launch_bounds__(256) void point_prep_kernel(PointArgs pa, const float2 *values,
                                                          Geom g, KbPoly poly, float4 *rec,
                                                          uint8_t *xo) {
  using R = XRec<W>;
  constexpr int RW = R::RW;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < pa.num_points; s += st) {
    const int64_t j = (int64_t)pa.perm[s];
    float r[RW];
    int off = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float *cp = a == 0 ? pa.x : (a == 1 ? pa.y : pa.z);
      int cell;
      const double t = fold_coord(cp[j * pa.stride], g.scale[a], g.m[a], g.inv_m[a], cell);
      kb_taps<W>(t, cell, poly, r + R::WX + a * 2 * W);
      const int o = cell - (cell / g.bin[a]) * g.bin[a];
      off |= (a == 0 ? o : o + 1) << (8 * a);
      if (a == 0) xo[s] = (uint8_t)o;
    }
    if (values) {
      const float2 c = values[j];
      r[0] = c.x;
      r[1] = c.y;
    } else {
      r[0] = 0.f;
      r[1] = 0.f;
    }
    r[R::OFF] = __int_as_float(off);
#pragma unroll
    for (int i = R::OFF + 1; i < RW; ++i) r[i] = 0.f;
    float4 *d = rec + s * (RW / 4);
#pragma unroll
    for (int q = 0; q < RW / 4; ++q) d[q] = make_float4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
  }
}

// Tap polynomials staged in LDS as rows of kPolyRow floats per tap (the
// kPolyCoefs coefficients, lowest degree first, zero padded), so that a lane
// evaluates all taps of a point with broadcast reads of uniform addresses.
constexpr int kPolyRow = 12;
__device__ __forceinline__ void stage_poly_rows(const float (*__restrict__ coef)[kMaxTaps], int T,
                                                float *rows) {
  for (int k = threadIdx.x; k < T * kPolyRow; k += blockDim.x) {
    const int tap = k / kPolyRow, d = k - tap * kPolyRow;
    rows[k] = d < kPolyCoefs ? coef[d][tap] : 0.f;
  }
}

// Weights of all 2W taps at two arguments z0, z1 (kb_taps's endpoint rule:
// the first tap vanishes at z <= -1, the last at z >= 1).
template <int W>
__device__ __forceinline__ void taps2(const float *__restrict__ rows, float z0, float z1,
                                      float (&v0)[2 * W], float (&v1)[2 * W]) {
#pragma unroll
  for (int i = 0; i < 2 * W; ++i) {
    float c[kPolyRow];
#pragma unroll
    for (int d = 0; d < kPolyRow; d += 4) *(float4 *)(c + d) = *(const float4 *)(rows + i * kPolyRow + d);
    float a = c[kPolyCoefs - 1], b = c[kPolyCoefs - 1];
#pragma unroll
    for (int d = kPolyCoefs - 2; d >= 0; --d) {
      a = fmaf(a, z0, c[d]);
      b = fmaf(b, z1, c[d]);
    }
    v0[i] = a;
    v1[i] = b;
  }
  if (z0 <= -1.0f) v0[0] = 0.f;
  if (z1 <= -1.0f) v1[0] = 0.f;
  if (z0 >= 1.0f) v0[2 * W - 1] = 0.f;
  if (z1 >= 1.0f) v1[2 * W - 1] = 0.f;
}

// taps2 from the mirrored tables (tap 2W-1-i at z is tap i at -z, negated
// for the derivative, ODD): each pair from the even part E(z^2) and odd part
// O(z^2) of tap i, P_i = E + z O, P_{2W-1-i} = E - z O (ODD: z O - E); the
// same endpoint rule.
static_assert(kPolyCoefs == 11, "taps2_sym splits degree-10 polynomials into even and odd parts");
template <int W, bool ODD = false>
__device__ __forceinline__ void taps2_sym(const float *__restrict__ rows, float z0, float z1,
                                          float (&v0)[2 * W], float (&v1)[2 * W]) {
  const float u0 = z0 * z0, u1 = z1 * z1;
#pragma unroll
  for (int i = 0; i < W; ++i) {
    float c[kPolyRow];
#pragma unroll
    for (int d = 0; d < kPolyRow; d += 4) *(float4 *)(c + d) = *(const float4 *)(rows + i * kPolyRow + d);
    float e0 = c[10], e1 = c[10], o0 = c[9], o1 = c[9];
#pragma unroll
    for (int d = 8; d >= 0; d -= 2) {
      e0 = fmaf(e0, u0, c[d]);
      e1 = fmaf(e1, u1, c[d]);
    }
#pragma unroll
    for (int d = 7; d >= 1; d -= 2) {
      o0 = fmaf(o0, u0, c[d]);
      o1 = fmaf(o1, u1, c[d]);
    }
    v0[i] = fmaf(z0, o0, e0);
    v1[i] = fmaf(z1, o1, e1);
    v0[2 * W - 1 - i] = ODD ? fmaf(z0, o0, -e0) : fmaf(-z0, o0, e0);
    v1[2 * W - 1 - i] = ODD ? fmaf(z1, o1, -e1) : fmaf(-z1, o1, e1);
  }
  if (z0 <= -1.0f) v0[0] = 0.f;
  if (z1 <= -1.0f) v1[0] = 0.f;
  if (z0 >= 1.0f) v0[2 * W - 1] = 0.f;
  if (z1 >= 1.0f) v1[2 * W - 1] = 0.f;
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
    const double t = fold_coord(cp[(int64_t)j * pa.stride], g.scale[ax], g.m[ax], g.inv_m[ax], cell);
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

// ---- MFMA spreading ---------------------------------------------------------
//
// The tile's (y,z) plane is a 32 x 32 MFMA output block per x cell: for x
// cell i, G_i[y][z] += sum_k A_i[y][k] B[k][z] with A_i[y][k] = wx_k(i) wy_k(y)
// and B[k][z] = c_k wz_k(z) (complex c: one block for re, one for im), k over
// the subproblem's points 16 at a time (v_mfma_f32_32x32x16_f16).  Each FP32
// operand is split into FP16 hi + lo (hi = fp16(a), lo = fp16(a - hi)) and a
// product is hi*hi + hi*lo + lo*hi, accumulated in FP32 by the MFMA: about
// 22 significant bits per product.  The strengths of a subproblem are scaled
// by a power of two so that max |c| is in [1/2, 1), which keeps every
// operand inside FP16's range; the flush undoes the scale exactly.  A wave
// owns XW = 4 consecutive x cells (4 x 2 accumulator blocks); the B
// operands of a point group are shared by the wave's cells, A is rebuilt per
// cell from the staged y and x weights.
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Operands enter the FP16 split multiplied by 2^12 (exact), so that the low
// halves of operands down to 2^-14 of the largest stay FP16-normal (11
// significant bits): values in [0, 1] become [0, 4096] (FP16 max 65504).
// Products carry 2^24, removed with the output scale.  Kernel derivatives
// (|phi'| < 64 for w <= 8) use 2^8.
constexpr float kOpScale = 4096.0f;
constexpr float kOpScaleD = 256.0f;
constexpr float kProdUnscale = 1.0f / (4096.0f * 4096.0f);
constexpr float kProdUnscaleD = 1.0f / (4096.0f * 256.0f);

template <int W>
struct MSpread {
  static constexpr int NW = W >= 8 ? 8 : 4;    // waves per workgroup
  static constexpr int XW = 4;                 // x cells per wave
  static constexpr int PX = NW * XW;           // x cells of a tile
  static constexpr int BINX = PX - 2 * W + 1;  // tile interior along x
  static constexpr int NT = 64 * NW;
};
constexpr int kMNP = 128;       // points staged per round
constexpr int kMWY = kMNP + 4;  // row stride (floats) of the y / x weight tables
constexpr int kMBT = kMNP + 8;  // row stride (halves) of the B operand tables
// LDS layout of the MFMA spreading kernels (bytes), for a tile of PX x cells.
constexpr int kMWZ = 33;        // row stride (floats) of the staged z weights
template <int PX>
struct MLds {
  static constexpr int WY = 0;                                 // float [32][kMWY]
  static constexpr int WX = WY + 32 * kMWY * 4;                // float [PX][kMWY]
  static constexpr int BT = WX + PX * kMWY * 4;                // half [4][32][kMBT]
  static constexpr int C = BT + 4 * 32 * kMBT * 2;             // float2 [kMNP]
  static constexpr int WZ = C + kMNP * 8;                      // float [kMNP][kMWZ]
  static constexpr int ZS = WZ + kMNP * kMWZ * 4;              // float [3][kMNP]
  static constexpr int OFS = ZS + 3 * kMNP * 4;                // int [kMNP]
  static constexpr int RED = OFS + kMNP * 4;                   // float [8]
  static constexpr int POLY = RED + 8 * 4;                     // float [kMaxTaps][kPolyRow]
  static constexpr int SIZE = POLY + kMaxTaps * kPolyRow * 4;
};
// two 4-wave workgroups per CU for w <= 7
static_assert(2 * MLds<16>::SIZE <= 160 * 1024, "MFMA spreading LDS");

typedef float float2v __attribute__((ext_vector_type(2)));
typedef _Float16 half2v __attribute__((ext_vector_type(2)));

// hi = fp16(a) (round to nearest, packed pairs), lo = fp16(a - hi) (round
// toward zero, packed pairs).  Both halves of a register are written by one
// packed conversion: split sequences that wrote the low and high halves of a
// register with separate half-register instructions gave sporadically wrong
// low parts on gfx950, while packed conversions of the same values did not.
__device__ __forceinline__ void split8(const float *a, half8 &hi, half8 &lo) {
#pragma unroll
  for (int j = 0; j < 8; j += 2) {
    const float2v v = {a[j], a[j + 1]};
    const half2v h = __builtin_convertvector(v, half2v);
    typedef __fp16 fp16x2 __attribute__((ext_vector_type(2)));
    const fp16x2 l = __builtin_amdgcn_cvt_pkrtz(a[j] - (float)h[0], a[j + 1] - (float)h[1]);
    hi[j] = h[0];
    hi[j + 1] = h[1];
    lo[j] = (_Float16)l[0];
    lo[j + 1] = (_Float16)l[1];
  }
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

// Maximum of non-negative values over the 64 lanes (the DPP steps of
// wave_sum; lanes a step does not write read 0, which is neutral here).
__device__ __forceinline__ float wave_max_nonneg(float v) {
  v = fmaxf(v, dpp<0xB1>(v));
  v = fmaxf(v, dpp<0x4E>(v));
  v = fmaxf(v, dpp<0x141>(v));
  v = fmaxf(v, dpp<0x140>(v));
  v = fmaxf(v, dpp<0x142, 0xa>(v));
  v = fmaxf(v, dpp<0x143, 0xc>(v));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}


// Tile of a spreading subproblem.  The accumulators of a workgroup form a
// ring of PX x columns: column x lives in slot x mod PX, so a workgroup that
// moves on to the next tile along x keeps the columns both tiles cover and
// flushes only those it leaves behind.
struct MTile {
  int pencil;       // y-z tile row (by * nbin_z + bz)
  int bx;           // x tile index
  int xlo, oy, oz;  // grid cell of tile column 0, row 0, layer 0
};

template <int W>
__device__ __forceinline__ MTile mfma_tile(const Geom &g, int bin) {
  const int bz = bin % g.nbin[2];
  const int by = (bin / g.nbin[2]) % g.nbin[1];
  const int bx = bin / (g.nbin[2] * g.nbin[1]);
  return MTile{by * g.nbin[2] + bz, bx, bx * g.bin[0] - W + 1, by * g.bin[1] - W,
               bz * g.bin[2] - W};
}

// Power-of-two scale of a subproblem's strengths: e with max |c| 2^-e in
// [1/2, 1), clamped to [-120, 120] so that both scales stay finite and normal.
template <int NW>
__device__ __forceinline__ int mfma_strength_exp(const ZRec &zr, int4 sub, float *red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int NT = 64 * NW, B = 4;  // B values per thread loaded before any is used
  const float2 *vals = (const float2 *)zr.r + 4 * (int64_t)sub.y + 2;  // point p's value: vals[4 p]
  float mx = 0.f;
  for (int p0 = threadIdx.x; p0 < sub.z; p0 += NT * B) {
    float2 c[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int p = p0 + u * NT;
      c[u] = vals[4 * (int64_t)(p < sub.z ? p : 0)];
    }
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (p0 + u * NT < sub.z) mx = fmaxf(mx, fmaxf(fabsf(c[u].x), fabsf(c[u].y)));
  }
  mx = wave_max_nonneg(mx);
  __syncthreads();  // earlier readers of red[] are done
  if (lane == 0) red[wave] = mx;
  __syncthreads();
  mx = red[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) mx = fmaxf(mx, red[i]);
  int e = 0;
  if (mx > 0.f) (void)frexpf(mx, &e);
  return max(-120, min(120, e));
}

// Add the points of one subproblem, strengths scaled by s_in, to the ring
// accumulators; base = tile column 0's ring slot.
template <int W>
__device__ __forceinline__ void mfma_accumulate(f32x16 (&acc)[MSpread<W>::XW][2], const ZRec &zr,
                                                const KbPoly *__restrict__ dpoly, int4 sub,
                                                int base, float s_in, char *sh) {
  using MS = MSpread<W>;
  constexpr int NT = MS::NT, XW = MS::XW, PX = MS::PX, T = 2 * W;
  using L = MLds<PX>;
  float *WY = (float *)(sh + L::WY);
  float *WX = (float *)(sh + L::WX);
  _Float16 *BT = (_Float16 *)(sh + L::BT);
  float2 *C = (float2 *)(sh + L::C);
  float *WZ = (float *)(sh + L::WZ);
  float *ZS = (float *)(sh + L::ZS);
  int *OFS = (int *)(sh + L::OFS);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int h = lane >> 5, l32 = lane & 31;

  for (int c0 = 0; c0 < sub.z; c0 += kMNP) {
    const int npts = min(kMNP, sub.z - c0);
    const int ng = (npts + 15) >> 4;
    const int np16 = ng << 4;
    const int start = sub.y + c0;
    __syncthreads();  // the previous round's operand reads are done
    // pass 0: zero the operand tables; the points' polynomial arguments,
    // tile offsets and scaled values into LDS (padding points: z = 0,
    // offsets 0, value 0, so their finite weights meet zero B rows)
    for (int p = threadIdx.x; p < np16; p += NT) {
#pragma unroll
      for (int i = 0; i < PX; ++i) WX[i * kMWY + p] = 0.f;
#pragma unroll
      for (int y = 0; y < 32; ++y) WY[y * kMWY + p] = 0.f;
#pragma unroll
      for (int zz = 0; zz < 32; ++zz) WZ[p * kMWZ + zz] = 0.f;
      const bool ok = p < npts;
      const float4 z = ok ? zr.r[2 * (int64_t)(start + p)] : make_float4(0.f, 0.f, 0.f, 0.f);
      const float4 c = ok ? zr.r[2 * (int64_t)(start + p) + 1] : make_float4(0.f, 0.f, 0.f, 0.f);
      ZS[p] = z.x;
      ZS[kMNP + p] = z.y;
      ZS[2 * kMNP + p] = z.z;
      OFS[p] = __float_as_int(z.w);
      C[p] = make_float2(c.x * s_in * kOpScale, c.y * s_in * kOpScale);
    }
    __syncthreads();
    // pass 1: tap weights; a task evaluates all taps of one axis for two
    // points (p, p + np16 / 2).  The x weights go to ring slot (tile column
    // + base) mod PX.
    {
      const float *PL = (const float *)(sh + L::POLY);
      const int hn = np16 >> 1;
      for (int u = threadIdx.x; u < 3 * hn; u += NT) {
        const int a = (u >= hn) + (u >= 2 * hn);
        const int p0 = u - a * hn, p1 = p0 + hn;
        const int o0 = (OFS[p0] >> (8 * a)) & 255, o1 = (OFS[p1] >> (8 * a)) & 255;
        float v0[T], v1[T];
        taps2_sym<W>(PL, ZS[a * kMNP + p0], ZS[a * kMNP + p1], v0, v1);
        if (a == 0) {
#pragma unroll
          for (int i = 0; i < T; ++i) {
            WX[((o0 + i + base) & (PX - 1)) * kMWY + p0] = v0[i] * kOpScale;
            WX[((o1 + i + base) & (PX - 1)) * kMWY + p1] = v1[i] * kOpScale;
          }
        } else if (a == 1) {
#pragma unroll
          for (int i = 0; i < T; ++i) {
            WY[(o0 + i) * kMWY + p0] = v0[i];
            WY[(o1 + i) * kMWY + p1] = v1[i];
          }
        } else {
#pragma unroll
          for (int i = 0; i < T; ++i) {
            WZ[p0 * kMWZ + o0 + i] = v0[i];
            WZ[p1 * kMWZ + o1 + i] = v1[i];
          }
        }
      }
    }
    __syncthreads();
    // pass 2: B operands c_k wz_k(z), split, 8 points per task
    for (int t = threadIdx.x; t < 32 * (np16 >> 3); t += NT) {
      const int z = t & 31, pb = (t >> 5) << 3;
      float br[8], bi[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float w = WZ[(pb + j) * kMWZ + z];
        const float2 c = C[pb + j];
        br[j] = c.x * w;
        bi[j] = c.y * w;
      }
      half8 rh, rl, ih, il;
      split8(br, rh, rl);
      split8(bi, ih, il);
      *(half8 *)(BT + (0 * 32 + z) * kMBT + pb) = rh;
      *(half8 *)(BT + (1 * 32 + z) * kMBT + pb) = rl;
      *(half8 *)(BT + (2 * 32 + z) * kMBT + pb) = ih;
      *(half8 *)(BT + (3 * 32 + z) * kMBT + pb) = il;
    }
    __syncthreads();
    // MFMA: 16 points per step; lane (l32, h) holds A row y = l32 and
    // B column z = l32 for the points kb .. kb + 7; wave w owns ring
    // slots w XW .. w XW + XW - 1
    for (int gi = 0; gi < ng; ++gi) {
      const int kb = (gi << 4) + (h << 3);
      const half8 brh = *(const half8 *)(BT + (0 * 32 + l32) * kMBT + kb);
      const half8 brl = *(const half8 *)(BT + (1 * 32 + l32) * kMBT + kb);
      const half8 bih = *(const half8 *)(BT + (2 * 32 + l32) * kMBT + kb);
      const half8 bil = *(const half8 *)(BT + (3 * 32 + l32) * kMBT + kb);
      float wy[8];
      *(float4 *)wy = *(const float4 *)(WY + l32 * kMWY + kb);
      *(float4 *)(wy + 4) = *(const float4 *)(WY + l32 * kMWY + kb + 4);
      half8 ah[XW], al[XW];
#pragma unroll
      for (int q = 0; q < XW; ++q) {
        const int ix = wave * XW + q;
        float a[8];
        *(float4 *)a = *(const float4 *)(WX + ix * kMWY + kb);
        *(float4 *)(a + 4) = *(const float4 *)(WX + ix * kMWY + kb + 4);
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] *= wy[j];
        split8(a, ah[q], al[q]);
      }
      // one split term at a time over all 2 XW accumulators: 2 XW
      // independent MFMA chains
#pragma unroll
      for (int q = 0; q < XW; ++q) {
        acc[q][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[q], brh, acc[q][0], 0, 0, 0);
        acc[q][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[q], bih, acc[q][1], 0, 0, 0);
      }
#pragma unroll
      for (int q = 0; q < XW; ++q) {
        acc[q][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[q], brl, acc[q][0], 0, 0, 0);
        acc[q][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[q], bil, acc[q][1], 0, 0, 0);
      }
#pragma unroll
      for (int q = 0; q < XW; ++q) {
        acc[q][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[q], brh, acc[q][0], 0, 0, 0);
        acc[q][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[q], bih, acc[q][1], 0, 0, 0);
      }
    }
  }
}

// Add the ring columns x in [t.xlo, t.xlo + ncol) to the grid (scaled by
// s_out) and zero them.  Accumulator r of lane (l32, h) is tile cell (y, z)
// with y = 8 (r / 4) + 4 h + r % 4, z = l32.
template <int W>
__device__ __forceinline__ void mfma_flush(f32x16 (&acc)[MSpread<W>::XW][2], float2 *grid,
                                           const Geom &g, const MTile &t, int ncol, float s_out) {
  using MS = MSpread<W>;
  constexpr int XW = MS::XW, PX = MS::PX;
  int lane = threadIdx.x & 63;
  // an opaque copy: keeps the per-row index arithmetic below from being
  // hoisted out of the subproblem loop, where it would stay live in
  // registers for the whole kernel
  asm volatile("" : "+v"(lane));
  const int wave = threadIdx.x >> 6;
  const int h = lane >> 5, l32 = lane & 31;
  int m0 = __builtin_amdgcn_readfirstlane(g.m[0]);
  int m1 = __builtin_amdgcn_readfirstlane(g.m[1]);
  int m2 = __builtin_amdgcn_readfirstlane(g.m[2]);
  asm volatile("" : "+s"(m0), "+s"(m1), "+s"(m2));  // no hoisted divisions by m
  const int base = t.xlo & (PX - 1);
  const bool zok = l32 < g.pad[2];
  const int gz = wrap(t.oz + l32, m2);
#pragma unroll
  for (int q = 0; q < XW; ++q) {
    const int ix = (wave * XW + q - base) & (PX - 1);  // tile column of the slot
    if (ix >= ncol) continue;
    if (zok) {
      const int gx = wrap(t.xlo + ix, m0);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int y = ((r >> 2) << 3) + (h << 2) + (r & 3);
        const float vr = acc[q][0][r] * s_out, vi = acc[q][1][r] * s_out;
        if (y < g.pad[1] && (vr != 0.f || vi != 0.f)) {
          const int gy = wrap(t.oy + y, m1);
          float2 *dst = grid + ((int64_t)gx * m1 + gy) * m2 + gz;
          unsafeAtomicAdd(&dst->x, vr);
          unsafeAtomicAdd(&dst->y, vi);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      acc[q][0][r] = 0.f;
      acc[q][1][r] = 0.f;
    }
  }
}

// Longest run of points accumulated in registers between full flushes.
constexpr int kMaxChain = 8192;

// Ring window of one workgroup: the open tile, its problem and scale.
struct MWindow {
  int prob = -1;  // problem of the open tile, -1 = none
  MTile t{};
  int e = 0;
};

// Columns of the open window to flush before the window moves to tile n of
// problem prob at strength exponent e (PX = all): a tile further along the
// same pencil keeps the columns both cover, unless the scales are so far
// apart that rescaling the kept sums could leave the normal FP32 range.
template <int W>
__device__ __forceinline__ int mfma_flush_cols(const MWindow &w, int prob, const MTile &n, int e,
                                               const Geom &g) {
  constexpr int PX = MSpread<W>::PX;
  if (w.prob != prob || w.t.pencil != n.pencil || n.bx < w.t.bx || abs(e - w.e) > 32) return PX;
  return min(PX, (n.bx - w.t.bx) * g.bin[0]);
}

template <int W>
__device__ __forceinline__ void mfma_rescale(f32x16 (&acc)[MSpread<W>::XW][2], float f) {
#pragma unroll
  for (int q = 0; q < MSpread<W>::XW; ++q) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      acc[q][0][r] *= f;
      acc[q][1][r] *= f;
    }
  }
}

// Work range [w0, w1) of this workgroup out of total work wt: each XCD
// takes one contiguous eighth (blocks are dealt to the XCDs round-robin,
// block b on XCD b mod 8) and splits it evenly among its blocks, so a
// workgroup walks consecutive tiles of a pencil and the XCD's tiles meet in
// its L2.
//
// Weights: blocks of the first half of the grid take wa shares, the others
// wb.  With two workgroups per CU the older one (dealt first) is issued
// ahead of the younger and runs faster, so equal shares leave the younger
// half finishing last; wa > wb hands the older half correspondingly more.
__device__ __forceinline__ void xcd_work_range(uint32_t wt, uint32_t &w0, uint32_t &w1,
                                               uint32_t wa = 1, uint32_t wb = 1) {
  constexpr uint32_t kXcd = 8;
  const uint32_t g = gridDim.x, b = blockIdx.x;
  if (g < kXcd) {
    w0 = (uint32_t)((uint64_t)wt * b / g);
    w1 = (uint32_t)((uint64_t)wt * (b + 1) / g);
    return;
  }
  const uint32_t x = b % kXcd, j = b / kXcd;
  const uint32_t per = (g - x + kXcd - 1) / kXcd;  // blocks on XCD x
  const uint32_t half = g / 2;                     // blocks j < nh of XCD x are in the first half
  const uint32_t nh = min(per, half > x ? (half - x + kXcd - 1) / kXcd : 0u);
  auto before = [&](uint32_t k) {  // shares of the XCD's blocks 0 .. k - 1
    const uint32_t m = min(k, nh);
    return (uint64_t)m * wa + (uint64_t)(k - m) * wb;
  };
  const uint64_t tot = before(per);
  const uint64_t lo = (uint64_t)wt * x / kXcd, hi = (uint64_t)wt * (x + 1) / kXcd;
  w0 = (uint32_t)(lo + (hi - lo) * before(j) / tot);
  w1 = (uint32_t)(lo + (hi - lo) * before(j + 1) / tot);
}

// First subproblem i in [0, n) whose work prefix subs[i].w is >= w (n if none).
__device__ __forceinline__ uint32_t work_lower_bound(const int4 *subs, uint32_t n, uint32_t w) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if ((uint32_t)subs[mid].w < w) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Shares of the older and younger workgroup of a CU in the interpolation
// launch when it holds two per CU (measured at 2e6 points, w = 7: the
// younger workgroups ran about 12 % longer on equal shares, 8 : 7 cut the
// launch by 4 %; spreading, bound by its matrix work per CU, did not
// respond to weights and keeps equal shares).
constexpr uint32_t kInterpOlder = 8, kInterpYounger = 7;

// The share of the subproblem list of this workgroup: the work range
// [w0, w1) of xcd_work_range over the list's work prefix W (W(i) =
// subs[i].w).  Subproblem i spans the work interval [W(i), W(i + 1)): its
// fixed cost first, then one unit per kWorkUnit points.  The workgroup owns the points whose units fall in
// its range, so the subproblems at the two ends of its run may be split with
// the neighbouring workgroups (each pays the fixed cost) and every
// workgroup gets the same work however coarse the subproblems are; every
// unit lies in exactly one range, so the shares partition the points.
// begin .. end - 1 are the subproblems whose intervals meet the range;
// points [h0, h1) of subproblem begin and [0, t1) of subproblem end - 1 are
// this workgroup's, every subproblem between is whole.
struct XcdChunk {
  uint32_t begin, end;
  int h0, h1, t1;
  __device__ __forceinline__ XcdChunk(const PointArgs &pa, uint32_t wa = 1, uint32_t wb = 1) {
    uint32_t w0, w1;
    xcd_work_range(pa.nsub[1], w0, w1, wa, wb);
    init(pa.subs, pa.nsub[0], pa.nsub[1], w0, w1);
  }
  // The share [w0, w1) of a list of n subproblems of total work wt.
  __device__ __forceinline__ XcdChunk(const int4 *subs, uint32_t n, uint32_t wt, uint32_t w0,
                                      uint32_t w1) {
    init(subs, n, wt, w0, w1);
  }
  __device__ __forceinline__ void init(const int4 *subs, uint32_t n, uint32_t wt, uint32_t w0,
                                       uint32_t w1) {
    const uint32_t b = work_lower_bound(subs, n, w0 + 1);  // first i with W(i) > w0
    begin = b > 0 ? b - 1 : 0;
    end = work_lower_bound(subs, n, w1);
    // drop a head or tail subproblem of which no point is this workgroup's
    // (the range starts past its points, or ends inside its fixed cost), so
    // that every subproblem of the run has points
    h0 = h1 = t1 = 0;
    int t0;
    if (begin < end) {
      owned(subs, n, wt, begin, w0, w1, h0, h1);
      if (h1 <= h0 && ++begin < end) owned(subs, n, wt, begin, w0, w1, h0, h1);
      if (h1 <= h0) begin = end;
    }
    if (end - begin > 1) {
      owned(subs, n, wt, end - 1, w0, w1, t0, t1);
      if (t1 <= 0 && --end - begin > 1) owned(subs, n, wt, end - 1, w0, w1, t0, t1);
    }
    begin = __builtin_amdgcn_readfirstlane(begin);
    end = __builtin_amdgcn_readfirstlane(end);
    h0 = __builtin_amdgcn_readfirstlane(h0);
    h1 = __builtin_amdgcn_readfirstlane(h1);
    t1 = __builtin_amdgcn_readfirstlane(t1);
  }
  // Points [p0, p1) of subproblem i whose units fall in [w0, w1).
  __device__ __forceinline__ static void owned(const int4 *subs, uint32_t n, uint32_t wt, uint32_t i,
                                               uint32_t w0, uint32_t w1, int &p0, int &p1) {
    const int4 sub = subs[i];
    const uint32_t wi = (uint32_t)sub.w, wn = i + 1 < n ? (uint32_t)subs[i + 1].w : wt;
    const uint32_t u = ((uint32_t)sub.z + kWorkUnit - 1) / kWorkUnit;
    const uint32_t a = wi + (wn - wi > u ? wn - wi - u : 0u);  // unit 0's place
    const uint32_t last = wn - 1;  // unit k sits at min(a + k, last)
    auto first = [&](uint32_t w) { return last < w ? u : (w <= a ? 0u : min(u, w - a)); };
    p0 = (int)(first(w0) * kWorkUnit);
    p1 = min(sub.z, (int)(first(w1) * kWorkUnit));
  }
  // Subproblem i (sub = subs[i]) of begin .. end - 1 narrowed to this
  // workgroup's points (at least one).
  __device__ __forceinline__ int4 part(uint32_t i, int4 sub) const {
    const int y = i == begin ? sub.y + h0 : sub.y;
    const int z = i == begin ? h1 - h0 : (i + 1 == end ? t1 : sub.z);
    // wave-uniform values: kept in scalar registers
    return make_int4(sub.x, __builtin_amdgcn_readfirstlane(y), __builtin_amdgcn_readfirstlane(z),
                     sub.w);
  }
};

// Spread the subproblems [begin, end) of one problem's list with one ring
// window: a tile further along the pencil keeps the columns both cover;
// everything left is flushed at the end.
template <int W, class Part>
__device__ __forceinline__ void spread_run_mfma(const int4 *__restrict__ subs, uint32_t begin,
                                                uint32_t end, const Part &part, const ZRec &zr,
                                                float2 *grid, const Geom &g,
                                                const KbPoly *__restrict__ dpoly, char *shm) {
  using MS = MSpread<W>;
  float *red = (float *)(shm + MLds<MS::PX>::RED);
  f32x16 acc[MS::XW][2];
#pragma unroll
  for (int q = 0; q < MS::XW; ++q)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[q][0][r] = acc[q][1][r] = 0.f;
  MWindow win;
  int chain = 0;  // points accumulated since the last full flush
  for (uint32_t sidx = begin; sidx < end; ++sidx) {
    const int4 sub = part.part(sidx, subs[sidx]);
    const MTile t = mfma_tile<W>(g, sub.x);
    const int e = mfma_strength_exp<MS::NW>(zr, sub, red);
    if (win.prob >= 0) {
      int ncol = mfma_flush_cols<W>(win, 0, t, e, g);
      // bound the FP32 accumulation chains (rounding error grows with their
      // length) as a full flush per subproblem of kMaxChain points would
      if (chain + sub.z > kMaxChain) ncol = MS::PX;
      mfma_flush<W>(acc, grid, g, win.t, ncol, ldexpf(1.0f, win.e) * kProdUnscale);
      if (ncol < MS::PX && e != win.e) mfma_rescale<W>(acc, ldexpf(1.0f, win.e - e));
      if (ncol == MS::PX) chain = 0;
    }
    chain += sub.z;
    win.prob = 0;
    win.t = t;
    win.e = e;
    mfma_accumulate<W>(acc, zr, dpoly, sub, t.xlo & (MS::PX - 1), ldexpf(1.0f, -e), shm);
  }
  if (win.prob >= 0)
    mfma_flush<W>(acc, grid, g, win.t, MS::PX, ldexpf(1.0f, win.e) * kProdUnscale);
}

template <int W>
__global__ __launch_bounds__(MSpread<W>::NT, 2) void spread3d_mfma_kernel(
    PointArgs pa, ZRec zr, float2 *grid, Geom g, const KbPoly *__restrict__ dpoly) {
  using MS = MSpread<W>;
  extern __shared__ __align__(16) char shm[];
  stage_poly_rows(dpoly->c, 2 * W, (float *)(shm + MLds<MS::PX>::POLY));  // ordered by the first barrier
  const XcdChunk r(pa);
  spread_run_mfma<W>(pa.subs, r.begin, r.end, r, zr, grid, g, dpoly, shm);
}

// Batched form: one persistent launch over the subproblems of all problems
// (prefix[p] = first global subproblem index of problem p, prefix[nprob] =
// total, then the work prefix of the problems).  A workgroup's work range
// maps to a run of global subproblems, which it spreads problem by problem.
template <int W>
__global__ __launch_bounds__(MSpread<W>::NT, 2) void spread3d_mfma_batched_kernel(
    const SpreadProblem *probs, const uint32_t *prefix, int nprob,
    const KbPoly *__restrict__ dpoly) {
  using MS = MSpread<W>;
  extern __shared__ __align__(16) char shm[];
  stage_poly_rows(dpoly->c, 2 * W, (float *)(shm + MLds<MS::PX>::POLY));  // ordered by the first barrier
  const uint32_t *wpre = prefix + nprob + 1;
  uint32_t w0, w1;
  xcd_work_range(wpre[nprob], w0, w1);
  // last problem with wpre <= w
  auto problem_of = [&](uint32_t w) {
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (wpre[mid] <= w) lo = mid;
      else hi = mid - 1;
    }
    return lo;
  };
  if (w0 >= w1) return;
  // each problem the range meets: its share of the problem's own list, as
  // for a single problem (subproblems at the range ends split)
  for (int p = problem_of(w0); p < nprob && wpre[p] < w1; ++p) {
    const uint32_t lo = wpre[p], hi = wpre[p + 1];
    if (hi <= w0) continue;  // no work of this problem in the range
    const SpreadProblem &pr = probs[p];
    const XcdChunk part(pr.pa.subs, prefix[p + 1] - prefix[p], hi - lo, max(w0, lo) - lo,
                        min(w1, hi) - lo);
    spread_run_mfma<W>(pr.pa.subs, part.begin, part.end, part, pr.zr, pr.grid, pr.g, dpoly, shm);
  }
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

// prefix[p] = sum of the subproblem counts of problems < p, and
// prefix[nprob + 1 + p] = sum of their work.
__global__ void batch_prefix_kernel(const SpreadProblem *probs, int nprob, uint32_t *prefix) {
  if (threadIdx.x != 0) return;
  uint32_t acc = 0, accw = 0;
  for (int p = 0; p < nprob; ++p) {
    prefix[p] = acc;
    prefix[nprob + 1 + p] = accw;
    acc += probs[p].pa.nsub[0];
    accw += probs[p].pa.nsub[1];
  }
  prefix[nprob] = acc;
  prefix[2 * nprob + 1] = accw;
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

// ---- tile maxima of the grid (scales of the MFMA interpolation tiles) -------
//
// bmax[b] = largest |re| or |im| over the cells of bin b; tmax[b] = largest
// bmax over the bins that the padded tile of bin b overlaps, i.e. an upper
// bound of the largest component of that tile.
//
// Bin maxima: one block per (x bin, y bin) column of bins; its rows (x, y)
// are read along z with a wave on 64 consecutive cells, each thread keeps
// the maximum of one z cell over its rows, then folds it into the z bin's
// slot in LDS (non-negative floats order like their bit patterns).
__global__ __launch_bounds__(256) void bin_max_kernel(const float2 *grid, Geom g, float *bmax) {
  extern __shared__ uint32_t bm[];  // nbin[2]
  const int bx = blockIdx.x / g.nbin[1], by = blockIdx.x - bx * g.nbin[1];
  for (int i = threadIdx.x; i < g.nbin[2]; i += 256) bm[i] = 0u;
  __syncthreads();
  const int x0 = bx * g.bin[0], y0 = by * g.bin[1];
  const int nx = min(g.bin[0], g.m[0] - x0), ny = min(g.bin[1], g.m[1] - y0);
  const int nrows = nx * ny, m2 = g.m[2];
  const int tz = threadIdx.x & 63, tr = threadIdx.x >> 6;
  for (int z = tz; z < m2; z += 64) {
    float mx = 0.f;
    for (int r = tr; r < nrows; r += 4) {
      const int x = x0 + r / ny, y = y0 + r % ny;
      const float2 v = grid[((int64_t)x * g.m[1] + y) * m2 + z];
      mx = fmaxf(mx, fmaxf(fabsf(v.x), fabsf(v.y)));
    }
    atomicMax(&bm[z / g.bin[2]], __float_as_uint(mx));
  }
  __syncthreads();
  float *dst = bmax + ((int64_t)bx * g.nbin[1] + by) * g.nbin[2];
  for (int i = threadIdx.x; i < g.nbin[2]; i += 256) dst[i] = __uint_as_float(bm[i]);
}

// Tile maxima: for the tile of bin b, the maximum of bmax over the bins that
// meet its cells [b * bin + org, + ext) along each axis (periodic).  One
// wave per tile: lanes 0-2 list the bins of one axis each in LDS, then the
// lanes share the combinations.
constexpr int kTileMaxList = 32;
__device__ __forceinline__ int axis_bins(int b, int bw, int m, int org, int ext, int *out) {
  int c = b * bw + org;
  while (c < 0) c += m;
  while (c >= m) c -= m;
  int n = 0, left = ext;
  while (left > 0 && n < kTileMaxList) {
    const int bin = c / bw;
    if (n == 0 || out[n - 1] != bin) out[n++] = bin;
    const int step = min(left, min((bin + 1) * bw, m) - c);  // to the next bin or the grid end
    left -= step;
    c += step;
    if (c >= m) c -= m;
  }
  return n;
}

__global__ __launch_bounds__(64) void tile_max_kernel(const float *bmax, Geom g, int3 org,
                                                      int3 ext, float *tmax) {
  __shared__ int lst[3][kTileMaxList];
  __shared__ int cnt[3];
  const int bin = blockIdx.x;
  const int b3[3] = {bin / (g.nbin[2] * g.nbin[1]), (bin / g.nbin[2]) % g.nbin[1], bin % g.nbin[2]};
  const int o3[3] = {org.x, org.y, org.z}, e3[3] = {ext.x, ext.y, ext.z};
  if (threadIdx.x < 3) {
    const int a = threadIdx.x;
    cnt[a] = axis_bins(b3[a], g.bin[a], g.m[a], o3[a], e3[a], lst[a]);
  }
  __syncthreads();
  const int nx = cnt[0], ny = cnt[1], nz = cnt[2];
  float mx = 0.f;
  for (int c = threadIdx.x; c < nx * ny * nz; c += 64) {
    const int k = c % nz, j = (c / nz) % ny, i = c / (nz * ny);
    mx = fmaxf(mx, bmax[((int64_t)lst[0][i] * g.nbin[1] + lst[1][j]) * g.nbin[2] + lst[2][k]]);
  }
  mx = wave_max(mx);
  if (threadIdx.x == 0) tmax[bin] = mx;
}

// ---- MFMA interpolation on 16-row windows ----------------------------------
//
// Same tiles as MFMA spreading; the tile is scaled by a power of two so that
// its largest component is in [1/2, 1) (from the tile maxima) and split into
// FP16 hi + lo like the spreading operands.  A point's 2w tap rows in y start
// at its tile offset o1 in [1, 32 - 2w] (the binning's record offset), so a
// 32 x 32 product block would use 2w of its 32 y rows.  Points whose o1 lie
// in one class of 17 - 2w consecutive offsets (one offset at w = 8) all fit
// one 16-row window of the tile, starting at row origin(c); for such a group
// of 16 points the y contraction runs on v_mfma_f32_16x16x32_f16 (16 window
// rows x 16 points, K = the tile's 32 z columns): half the matrix work per
// point of the 32 x 32 form.
//
// The tile lives in registers as two 16-row halves: lane (i, g) (i = lane %
// 16, g = lane / 16) holds rows i and 16 + i of the z columns 8g .. 8g + 7,
// which is the A-operand layout of the 16 x 16 x 32 product for the window at
// row 0 or 16.  The window of a class is rotated out of the halves once per
// class: lane i' of a 16-lane row passes its lower row if i' >= origin, else
// its upper row, and lane i takes lane (i + origin) mod 16's value.
//
// Each piece of at most kJPiece points of a subproblem is listed in LDS by
// class, each class padded to whole groups of 16 (pad slots compute on zero
// arguments and write nothing).  The order inside a class does not matter:
// each point's sum is formed in a fixed order from its own weights.
template <int W>
struct MWin {
  static constexpr int CW = 17 - 2 * W;                    // offsets per class
  static constexpr int NCLS = (32 - 2 * W + CW - 1) / CW;  // classes of o1 in [1, 32 - 2w]
  __device__ static int origin(int c) { return min(1 + c * CW, 16); }
  __device__ static int cls(int o1) { return min((o1 - 1) / CW, NCLS - 1); }
};
constexpr int kJNP = 128;      // slots staged per round (8 groups of 16)
constexpr int kJRowB = 40;     // row stride (halves) of the staged B operands (32 z columns)
constexpr int kJRowY = 20;     // row stride (floats) of the y weights (16 window rows)
constexpr int kJRowX = 20;     // of the x weights (16 columns)
constexpr int kJPiece = 4096;  // points listed by class at once
constexpr int kMaxCls = 16;    // classes at most (w = 8: one per offset)
constexpr int kClsInts = 3 * kMaxCls + 4;  // counts, cursors, starts (kMaxCls + 1), padding
constexpr int kListLen = kJPiece + 16 * kMaxCls;  // list entries with each class's padding
template <int PX, int NW>
struct JLds {
  static constexpr int ROWX = PX + 4;                   // row stride (floats) of the x weights
  static constexpr int BH = 0;                          // half [kJNP][kJRowB]: hi of z weights
  static constexpr int BL = BH + kJNP * kJRowB * 2;     // half [kJNP][kJRowB]: lo
  static constexpr int WY = BL + kJNP * kJRowB * 2;     // float [kJNP][kJRowY]
  static constexpr int WX = WY + kJNP * kJRowY * 4;     // float [kJNP][ROWX]
  static constexpr int RED = WX + kJNP * ROWX * 4;      // float2 [4 NW][kJNP]: per wave, lane row
  static constexpr int JIDX = RED + 4 * NW * kJNP * 8;  // int [kJNP]
  static constexpr int ZS = JIDX + kJNP * 4;            // float [3][kJNP]
  static constexpr int OFS = ZS + 3 * kJNP * 4;         // int [kJNP]
  static constexpr int CNT = OFS + kJNP * 4;            // int [kClsInts]
  static constexpr int LIST = CNT + kClsInts * 4;       // u16 [kListLen]
  static constexpr int CLS = LIST + kListLen * 2;       // u8 [kJPiece]: class of each point
  static constexpr int POLY = CLS + kJPiece;            // float [kMaxTaps][kPolyRow]
  static constexpr int SIZE = POLY + kMaxTaps * kPolyRow * 4;
};
static_assert(2 * JLds<16, 4>::SIZE <= 160 * 1024, "windowed interpolation: two workgroups per CU");
static_assert(JLds<32, 8>::SIZE <= 160 * 1024, "windowed interpolation at w = 8: one workgroup per CU");
static_assert(kJPiece < 0xFFFF, "list entries are u16 with 0xFFFF for padding");

typedef int int4v __attribute__((ext_vector_type(4)));

// Window rows origin .. origin + 15 from the two 16-row halves (see above).
__device__ __forceinline__ half8 window_rows(const half8 &lower, const half8 &upper, int origin,
                                             int i16, int addr) {
  const int4v l = __builtin_bit_cast(int4v, lower), u = __builtin_bit_cast(int4v, upper);
  const bool low = i16 >= origin;
  int4v r;
#pragma unroll
  for (int d = 0; d < 4; ++d) r[d] = __builtin_amdgcn_ds_bpermute(addr, low ? l[d] : u[d]);
  return __builtin_bit_cast(half8, r);
}

// Workgroups per CU of the windowed interpolation: two 4-wave groups for
// w <= 7 (16 x cells), one 8-wave group at w = 8 (32 x cells); 256
// registers per wave either way.
// The class list of one piece of np points (records base .. base + np - 1):
// the count of each class (one ballot per class and wave), class starts
// padded to whole groups of 16, each point's local index at its class's
// cursor and 0xFFFF in the pad slots.  The classes come from the records'
// offset words, loaded kClsBatch per thread before any is used, and are kept
// as bytes in CLS for the second pass.  Ends with a barrier.
constexpr int kClsBatch = 4;
template <class MW, int NT>
__device__ __forceinline__ void class_list(const ZRec &zr, int64_t base, int np, int *CNT,
                                           uint8_t *CLS, uint16_t *LIST) {
  constexpr int NCLS = MW::NCLS;
  int *CUR = CNT + kMaxCls;
  int *START = CNT + 2 * kMaxCls;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float *ofs = (const float *)zr.r + 8 * base + 3;  // record p's offset word: ofs[8 p]
  if (threadIdx.x < NCLS) CNT[threadIdx.x] = 0;
  __syncthreads();
  for (int pw = wave * 64; pw < np; pw += NT * kClsBatch) {
    int c[kClsBatch];
#pragma unroll
    for (int u = 0; u < kClsBatch; ++u) {
      const int p = pw + u * NT + lane;
      const int o = __float_as_int(ofs[8 * (int64_t)(p < np ? p : 0)]);
      c[u] = p < np ? MW::cls((o >> 8) & 255) : -1;
    }
#pragma unroll
    for (int u = 0; u < kClsBatch; ++u) {
      if (pw + u * NT >= np) break;  // wave-uniform
      const int p = pw + u * NT + lane;
      if (p < np) CLS[p] = (uint8_t)c[u];
#pragma unroll
      for (int k = 0; k < NCLS; ++k) {
        const uint64_t m = __ballot(c[u] == k);
        if (lane == 0 && m) atomicAdd(&CNT[k], (int)__popcll(m));
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int k = 0; k < NCLS; ++k) {
      START[k] = s;
      CUR[k] = s;
      s += (CNT[k] + 15) & ~15;
    }
    START[NCLS] = s;
  }
  __syncthreads();
  for (int pw = wave * 64; pw < np; pw += NT) {
    const int p = pw + lane;
    const int c = p < np ? (int)CLS[p] : -1;
    int pos = 0;
#pragma unroll
    for (int k = 0; k < NCLS; ++k) {
      const uint64_t m = __ballot(c == k);
      if (m) {
        int b = 0;
        if (lane == 0) b = atomicAdd(&CUR[k], (int)__popcll(m));
        b = __builtin_amdgcn_readfirstlane(b);
        const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (c == k) pos = b + rank;
      }
    }
    if (c >= 0) LIST[pos] = (uint16_t)p;
  }
  for (int t = threadIdx.x; t < 16 * NCLS; t += NT) {
    const int k = t >> 4;
    const int pos = START[k] + CNT[k] + (t & 15);
    if (pos < START[k + 1]) LIST[pos] = 0xFFFF;
  }
  __syncthreads();
}

template <int W>
constexpr int win_groups_per_cu() { return W <= 7 ? 2 : 1; }

template <int W>
__global__ __launch_bounds__(MSpread<W>::NT, win_groups_per_cu<W>()) void interp3d_win_kernel(
    PointArgs pa, ZRec zr, const float2 *grid, const float *tmax, float2 *out, Geom g,
    const KbPoly *__restrict__ dpoly, uint32_t wa, uint32_t wb) {
  using MS = MSpread<W>;
  using MW = MWin<W>;
  constexpr int NT = MS::NT, XW = MS::XW, NW = MS::NW, PX = MS::PX, NCLS = MW::NCLS;
  static_assert(XW == 4 && PX == 4 * NW && NCLS <= kMaxCls, "windowed interpolation shape");
  using L = JLds<PX, NW>;
  constexpr int ROWX = L::ROWX;
  extern __shared__ __align__(16) char shm[];
  _Float16 *BH = (_Float16 *)(shm + L::BH);
  _Float16 *BL = (_Float16 *)(shm + L::BL);
  float *WY = (float *)(shm + L::WY);
  float *WX = (float *)(shm + L::WX);
  float2 *RED = (float2 *)(shm + L::RED);
  int *JIDX = (int *)(shm + L::JIDX);
  float *ZS = (float *)(shm + L::ZS);
  int *OFS = (int *)(shm + L::OFS);
  int *CNT = (int *)(shm + L::CNT);
  int *START = CNT + 2 * kMaxCls;
  uint16_t *LIST = (uint16_t *)(shm + L::LIST);
  uint8_t *CLS = (uint8_t *)(shm + L::CLS);
  float *PL = (float *)(shm + L::POLY);
  stage_poly_rows(dpoly->c, 2 * W, PL);  // ordered by the first piece's barriers

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i16 = lane & 15, g4 = lane >> 4;
  const XcdChunk xr(pa, wa, wb);
  for (uint32_t sidx = xr.begin; sidx < xr.end; ++sidx) {
    const int4 sub = xr.part(sidx, pa.subs[sidx]);
    const int bin = sub.x;
    const int bz = bin % g.nbin[2];
    const int by = (bin / g.nbin[2]) % g.nbin[1];
    const int bx = bin / (g.nbin[2] * g.nbin[1]);
    const int bx0 = bx * g.bin[0];
    const int oy = by * g.bin[1] - W, oz = bz * g.bin[2] - W;
    const float mx = tmax[bin];
    int e = 0;
    if (mx > 0.f) (void)frexpf(mx, &e);
    e = max(-120, min(120, e));
    const float s_in = ldexpf(1.0f, -e) * kOpScale, s_out = ldexpf(1.0f, e) * kProdUnscale;

    // the tile: G[x][y = 16 hf + i16][z = 8 g4 + j] of this wave's XW x cells
    half8 trh[XW][2], trl[XW][2], tih[XW][2], til[XW][2];
    const int zc = 8 * g4, z0 = oz + zc;
#pragma unroll
    for (int q = 0; q < XW; ++q) {
      const int gx = wrap(bx0 - W + 1 + wave * XW + q, g.m[0]);
#pragma unroll
      for (int hf = 0; hf < 2; ++hf) {
        const int ty = 16 * hf + i16;
        const bool yok = ty < g.pad[1];
        const int gy = wrap(oy + ty, g.m[1]);
        const float2 *rowp = grid + ((int64_t)gx * g.m[1] + gy) * g.m[2];
        const int64_t e0 = ((int64_t)gx * g.m[1] + gy) * g.m[2] + z0;  // element index
        float2 v[8];
        if (yok && zc + 8 <= g.pad[2] && z0 >= 0 && z0 + 8 <= g.m[2] && (e0 & 1) == 0) {
          const float4 *src = (const float4 *)(rowp + z0);
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const float4 t = src[j];
            v[2 * j] = make_float2(t.x, t.y);
            v[2 * j + 1] = make_float2(t.z, t.w);
          }
        } else {
#pragma unroll
          for (int j = 0; j < 8; ++j)
            v[j] = (yok && zc + j < g.pad[2]) ? rowp[wrap(z0 + j, g.m[2])] : make_float2(0.f, 0.f);
        }
        float tr[8], ti[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          tr[j] = v[j].x * s_in;
          ti[j] = v[j].y * s_in;
        }
        split8(tr, trh[q][hf], trl[q][hf]);
        split8(ti, tih[q][hf], til[q][hf]);
      }
    }

    half8 wrh[XW], wrl[XW], wih[XW], wil[XW];  // the current class's window
    int wcls = -1;
    for (int p0 = 0; p0 < sub.z; p0 += kJPiece) {
      const int np = min(kJPiece, sub.z - p0);
      const int64_t base = (int64_t)sub.y + p0;
      __syncthreads();  // the previous piece's list, rows and sums are consumed
      class_list<MW, NT>(zr, base, np, CNT, CLS, LIST);
      int cstart[NCLS + 1];
#pragma unroll
      for (int k = 0; k <= NCLS; ++k) cstart[k] = __builtin_amdgcn_readfirstlane(START[k]);
      const int total = cstart[NCLS];

      for (int r0 = 0; r0 < total; r0 += kJNP) {
        const int nslot = min(kJNP, total - r0);  // a multiple of 16
        if (r0 > 0) __syncthreads();  // the previous round's rows and sums are consumed
        // stage: slot s holds list entry r0 + s; the y offset becomes relative
        // to the window of the slot's class
        for (int s = threadIdx.x; s < nslot; s += NT) {
          const int li = LIST[r0 + s];
          const bool ok = li != 0xFFFF;
          const int gpos = r0 + (s & ~15);
          int c = 0;
#pragma unroll
          for (int k = 1; k < NCLS; ++k) c += gpos >= cstart[k];
#pragma unroll
          for (int i = 0; i < 32; i += 8) {
            *(float4 *)(BH + s * kJRowB + i) = make_float4(0.f, 0.f, 0.f, 0.f);
            *(float4 *)(BL + s * kJRowB + i) = make_float4(0.f, 0.f, 0.f, 0.f);
          }
#pragma unroll
          for (int i = 0; i < 16; i += 4) *(float4 *)(WY + s * kJRowY + i) = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
          for (int i = 0; i < PX; i += 4) *(float4 *)(WX + s * ROWX + i) = make_float4(0.f, 0.f, 0.f, 0.f);
          const float4 z = ok ? zr.r[2 * (base + li)] : make_float4(0.f, 0.f, 0.f, 0.f);
          ZS[s] = z.x;
          ZS[kJNP + s] = z.y;
          ZS[2 * kJNP + s] = z.z;
          OFS[s] = ok ? __float_as_int(z.w) - (MW::origin(c) << 8) : 0;
          JIDX[s] = ok ? (int)__float_as_uint(zr.r[2 * (base + li) + 1].w) : -1;
        }
        __syncthreads();
        // tap weights: a task evaluates all taps of one axis for slots s, s + nslot / 2;
        // the z weights go to LDS as the B operands (scaled, FP16 hi + lo as split8)
        {
          const int hn = nslot >> 1;
          for (int u = threadIdx.x; u < 3 * hn; u += NT) {
            const int a = (u >= hn) + (u >= 2 * hn);
            const int s0 = u - a * hn, s1 = s0 + hn;
            const int o0 = (OFS[s0] >> (8 * a)) & 255, o1 = (OFS[s1] >> (8 * a)) & 255;
            float v0[2 * W], v1[2 * W];
            taps2<W>(PL, ZS[a * kJNP + s0], ZS[a * kJNP + s1], v0, v1);
            if (a == 2) {
#pragma unroll
              for (int t = 0; t < 2 * W; ++t) {
                const float2v x = {v0[t] * kOpScale, v1[t] * kOpScale};
                const half2v h = __builtin_convertvector(x, half2v);
                typedef __fp16 fp16x2 __attribute__((ext_vector_type(2)));
                const fp16x2 l = __builtin_amdgcn_cvt_pkrtz(x[0] - (float)h[0], x[1] - (float)h[1]);
                BH[s0 * kJRowB + o0 + t] = h[0];
                BH[s1 * kJRowB + o1 + t] = h[1];
                BL[s0 * kJRowB + o0 + t] = (_Float16)l[0];
                BL[s1 * kJRowB + o1 + t] = (_Float16)l[1];
              }
            } else {
              float *tab = a == 0 ? WX : WY;
              const int stride = a == 0 ? ROWX : kJRowY;
#pragma unroll
              for (int t = 0; t < 2 * W; ++t) {
                tab[s0 * stride + o0 + t] = v0[t];
                tab[s1 * stride + o1 + t] = v1[t];
              }
            }
          }
        }
        __syncthreads();
        for (int gi = 0; gi < (nslot >> 4); ++gi) {
          const int gpos = r0 + 16 * gi;
          int c = 0;
#pragma unroll
          for (int k = 1; k < NCLS; ++k) c += gpos >= cstart[k];
          if (c != wcls) {
            wcls = c;
            const int o = MW::origin(c);
            const int addr = 4 * (16 * g4 + ((i16 + o) & 15));
#pragma unroll
            for (int q = 0; q < XW; ++q) {
              wrh[q] = window_rows(trh[q][0], trh[q][1], o, i16, addr);
              wrl[q] = window_rows(trl[q][0], trl[q][1], o, i16, addr);
              wih[q] = window_rows(tih[q][0], tih[q][1], o, i16, addr);
              wil[q] = window_rows(til[q][0], til[q][1], o, i16, addr);
            }
          }
          const int k = 16 * gi + i16;  // this lane's slot
          const half8 bh = *(const half8 *)(BH + k * kJRowB + 8 * g4);
          const half8 bl = *(const half8 *)(BL + k * kJRowB + 8 * g4);
          // wy at this lane's D rows (window rows 4 g4 .. 4 g4 + 3), wx at the wave's cells
          const float4 wy = *(const float4 *)(WY + k * kJRowY + 4 * g4);
          const float4 wx4 = *(const float4 *)(WX + k * ROWX + wave * XW);
          const float wx[4] = {wx4.x, wx4.y, wx4.z, wx4.w};
          float accr = 0.f, acci = 0.f;
          // products of cell q + 1 are issued before the contraction of cell q
          const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
          f32x4 dr[2], di[2];
#pragma unroll
          for (int q = 0; q <= XW; ++q) {
            if (q < XW) {
              f32x4 &r = dr[q & 1], &i = di[q & 1];
              r = __builtin_amdgcn_mfma_f32_16x16x32_f16(wrh[q], bh, zero, 0, 0, 0);
              i = __builtin_amdgcn_mfma_f32_16x16x32_f16(wih[q], bh, zero, 0, 0, 0);
              r = __builtin_amdgcn_mfma_f32_16x16x32_f16(wrh[q], bl, r, 0, 0, 0);
              i = __builtin_amdgcn_mfma_f32_16x16x32_f16(wih[q], bl, i, 0, 0, 0);
              r = __builtin_amdgcn_mfma_f32_16x16x32_f16(wrl[q], bh, r, 0, 0, 0);
              i = __builtin_amdgcn_mfma_f32_16x16x32_f16(wil[q], bh, i, 0, 0, 0);
            }
            if (q > 0) {
              const f32x4 &r = dr[(q - 1) & 1], &i = di[(q - 1) & 1];
              const float pr = fmaf(wy.x, r[0], fmaf(wy.y, r[1], fmaf(wy.z, r[2], wy.w * r[3])));
              const float pi = fmaf(wy.x, i[0], fmaf(wy.y, i[1], fmaf(wy.z, i[2], wy.w * i[3])));
              accr = fmaf(wx[q - 1], pr, accr);
              acci = fmaf(wx[q - 1], pi, acci);
            }
          }
          // the four lane rows' partial sums are added in the output pass
          RED[(wave * 4 + g4) * kJNP + k] = make_float2(accr, acci);
        }
        __syncthreads();
        for (int s = threadIdx.x; s < nslot; s += NT) {
          const int j = JIDX[s];
          if (j < 0) continue;
          float2 v = RED[s];
#pragma unroll
          for (int r = 1; r < 4 * NW; ++r) {
            const float2 u = RED[r * kJNP + s];
            v.x += u.x;
            v.y += u.y;
          }
          out[j] = make_float2(v.x * s_out, v.y * s_out);
        }
      }
    }
  }
}

// ---- MFMA coordinate gradient --------------------------------------------
//
// With D_i[y][k] = sum_z G_i[y][z] wz_k(z) and E_i[y][k] = sum_z G_i[y][z]
// wz'_k(z) (two MFMA products per x cell), the three derivative sums are
//   Sx_k = sum_i wx'_k(i) sum_y wy_k(y) D_i,  Sy_k = sum_i wx_k(i) sum_y wy'_k(y) D_i,
//   Sz_k = sum_i wx_k(i) sum_y wy_k(y) E_i,
// and grad_a = -scale_a Re(conj(u_k) S_a), u_k read from the point's record
// (the binning stores u in the value fields).  Two x cells per wave (the
// second B operand and the derivative rows need the registers).  The
// products run on the 16-row windows of the windowed interpolation kernel.
template <int W>
struct MGrad {
  static constexpr int PX = MSpread<W>::PX;
  static constexpr int XW = 2;
  static constexpr int NW = PX / XW;
  static constexpr int NT = 64 * NW;
};
// ---- Windowed MFMA coordinate gradient (2w <= 14) --------------------------
//
// The sums of the gradient kernel above on the 16-row windows of the windowed
// interpolation kernel: for a group of 16 points of one y-offset class and x
// cell i of the wave, D_i = window x wz and E_i = window x wz' on
// v_mfma_f32_16x16x32_f16 (B = the points' z weights and their derivatives,
// staged as FP16 hi + lo rows), then Sx += wx'_i (wy . D_i), Sy += wx_i
// (wy' . D_i) and Sz += wx_i (wy . E_i) over this lane's four D rows.  Eight
// waves of two x cells each, one workgroup per CU; the tile maximum is taken
// from the tile itself.
constexpr int kGJNP = 128;  // slots staged per round (8 groups of 16)
template <int NW>
struct GJLds {
  static constexpr int BH = 0;                             // half [kGJNP][kJRowB]: wz hi
  static constexpr int BL = BH + kGJNP * kJRowB * 2;       // wz lo
  static constexpr int DH = BL + kGJNP * kJRowB * 2;       // wz' hi
  static constexpr int DL = DH + kGJNP * kJRowB * 2;       // wz' lo
  static constexpr int WY = DL + kGJNP * kJRowB * 2;       // float [kGJNP][kJRowY]: wy
  static constexpr int DY = WY + kGJNP * kJRowY * 4;       // wy'
  static constexpr int WX = DY + kGJNP * kJRowY * 4;       // float [kGJNP][kJRowX]: wx
  static constexpr int DX = WX + kGJNP * kJRowX * 4;       // wx'
  static constexpr int RED = DX + kGJNP * kJRowX * 4;      // float [NW][2][6][kGJNP]
  static constexpr int JIDX = RED + NW * 2 * 6 * kGJNP * 4;  // int [kGJNP]
  static constexpr int U = JIDX + kGJNP * 4;               // float2 [kGJNP]
  static constexpr int ZS = U + kGJNP * 8;                 // float [3][kGJNP]
  static constexpr int OFS = ZS + 3 * kGJNP * 4;           // int [kGJNP]
  static constexpr int TMX = OFS + kGJNP * 4;              // float [16]
  static constexpr int CNT = TMX + 16 * 4;                 // int [kClsInts]
  static constexpr int LIST = CNT + kClsInts * 4;          // u16 [kListLen]
  static constexpr int CLS = LIST + kListLen * 2;          // u8 [kJPiece]: class of each point
  static constexpr int POLY = CLS + kJPiece;               // float [2][kMaxTaps][kPolyRow]
  static constexpr int SIZE = POLY + 2 * kMaxTaps * kPolyRow * 4;
  static_assert(SIZE <= 160 * 1024, "one windowed gradient workgroup per CU");
};

// FP16 hi + lo of s v0 and s v1 (split8's rounding) at row offsets o0 + t, o1 + t.
template <int T>
__device__ __forceinline__ void store_split_rows(_Float16 *hi, _Float16 *lo, int r0, int r1,
                                                 const float (&v0)[T], const float (&v1)[T],
                                                 float s) {
#pragma unroll
  for (int t = 0; t < T; ++t) {
    const float2v x = {v0[t] * s, v1[t] * s};
    const half2v h = __builtin_convertvector(x, half2v);
    typedef __fp16 fp16x2 __attribute__((ext_vector_type(2)));
    const fp16x2 l = __builtin_amdgcn_cvt_pkrtz(x[0] - (float)h[0], x[1] - (float)h[1]);
    hi[r0 + t] = h[0];
    hi[r1 + t] = h[1];
    lo[r0 + t] = (_Float16)l[0];
    lo[r1 + t] = (_Float16)l[1];
  }
}

template <int W>
__global__ __launch_bounds__(MGrad<W>::NT, 1) void grad3d_win_kernel(
    PointArgs pa, ZRec zr, const float2 *grid, float *gx, float *gy, float *gz, Geom g,
    const KbPoly *__restrict__ dpoly) {
  using MG = MGrad<W>;
  using MW = MWin<W>;
  constexpr int NT = MG::NT, XW = MG::XW, NW = MG::NW, NCLS = MW::NCLS;
  static_assert(XW == 2 && MG::PX == 16 && NCLS <= kMaxCls, "windowed gradient shape");
  using L = GJLds<NW>;
  extern __shared__ __align__(16) char shm[];
  _Float16 *BH = (_Float16 *)(shm + L::BH);
  _Float16 *BL = (_Float16 *)(shm + L::BL);
  _Float16 *DH = (_Float16 *)(shm + L::DH);
  _Float16 *DL = (_Float16 *)(shm + L::DL);
  float *WY = (float *)(shm + L::WY);
  float *DY = (float *)(shm + L::DY);
  float *WX = (float *)(shm + L::WX);
  float *DX = (float *)(shm + L::DX);
  float *RED = (float *)(shm + L::RED);
  int *JIDX = (int *)(shm + L::JIDX);
  float2 *U = (float2 *)(shm + L::U);
  float *ZS = (float *)(shm + L::ZS);
  int *OFS = (int *)(shm + L::OFS);
  float *tmx = (float *)(shm + L::TMX);
  int *CNT = (int *)(shm + L::CNT);
  int *START = CNT + 2 * kMaxCls;
  uint16_t *LIST = (uint16_t *)(shm + L::LIST);
  uint8_t *CLS = (uint8_t *)(shm + L::CLS);
  float *PL = (float *)(shm + L::POLY);
  float *PLD = PL + kMaxTaps * kPolyRow;
  stage_poly_rows(dpoly->c, 2 * W, PL);  // ordered by the first subproblem's barriers
  stage_poly_rows(dpoly->dc, 2 * W, PLD);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i16 = lane & 15, g4 = lane >> 4;
  const XcdChunk xr(pa);
  for (uint32_t sidx = xr.begin; sidx < xr.end; ++sidx) {
    const int4 sub = xr.part(sidx, pa.subs[sidx]);
    const int bin = sub.x;
    const int bz = bin % g.nbin[2];
    const int by = (bin / g.nbin[2]) % g.nbin[1];
    const int bx = bin / (g.nbin[2] * g.nbin[1]);
    const int bx0 = bx * g.bin[0];
    const int oy = by * g.bin[1] - W, oz = bz * g.bin[2] - W;

    // the tile: G[x][y = 16 hf + i16][z = 8 g4 + j] of this wave's XW x cells,
    // scaled by the power of two of the workgroup's largest component
    const int zc = 8 * g4;
    float2 v[XW][2][8];
#pragma unroll
    for (int q = 0; q < XW; ++q) {
      const int gxx = wrap(bx0 - W + 1 + wave * XW + q, g.m[0]);
#pragma unroll
      for (int hf = 0; hf < 2; ++hf) {
        const int ty = 16 * hf + i16;
        const bool yok = ty < g.pad[1];
        const float2 *rowp = grid + ((int64_t)gxx * g.m[1] + wrap(oy + ty, g.m[1])) * g.m[2];
#pragma unroll
        for (int j = 0; j < 8; ++j)
          v[q][hf][j] = (yok && zc + j < g.pad[2]) ? rowp[wrap(oz + zc + j, g.m[2])]
                                                    : make_float2(0.f, 0.f);
      }
    }
    float mx = 0.f;
#pragma unroll
    for (int q = 0; q < XW; ++q)
#pragma unroll
      for (int hf = 0; hf < 2; ++hf)
#pragma unroll
        for (int j = 0; j < 8; ++j) mx = fmaxf(mx, fmaxf(fabsf(v[q][hf][j].x), fabsf(v[q][hf][j].y)));
    mx = wave_max(mx);
    __syncthreads();  // the previous subproblem's readers of tmx, rows and sums are done
    if (lane == 0) tmx[wave] = mx;
    __syncthreads();
    mx = tmx[0];
#pragma unroll
    for (int i = 1; i < NW; ++i) mx = fmaxf(mx, tmx[i]);
    int e = 0;
    if (mx > 0.f) (void)frexpf(mx, &e);
    e = max(-120, min(120, e));
    const float s_in = ldexpf(1.0f, -e) * kOpScale, s_out = ldexpf(1.0f, e);
    half8 trh[XW][2], trl[XW][2], tih[XW][2], til[XW][2];
#pragma unroll
    for (int q = 0; q < XW; ++q)
#pragma unroll
      for (int hf = 0; hf < 2; ++hf) {
        float tr[8], ti[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          tr[j] = v[q][hf][j].x * s_in;
          ti[j] = v[q][hf][j].y * s_in;
        }
        split8(tr, trh[q][hf], trl[q][hf]);
        split8(ti, tih[q][hf], til[q][hf]);
      }

    half8 wrh[XW], wrl[XW], wih[XW], wil[XW];  // the current class's window
    int wcls = -1;
    for (int p0 = 0; p0 < sub.z; p0 += kJPiece) {
      const int np = min(kJPiece, sub.z - p0);
      const int64_t base = (int64_t)sub.y + p0;
      if (p0 > 0) __syncthreads();  // the previous piece's list, rows and sums are consumed
      class_list<MW, NT>(zr, base, np, CNT, CLS, LIST);
      int cstart[NCLS + 1];
#pragma unroll
      for (int k = 0; k <= NCLS; ++k) cstart[k] = __builtin_amdgcn_readfirstlane(START[k]);
      const int total = cstart[NCLS];

      for (int r0 = 0; r0 < total; r0 += kGJNP) {
        const int nslot = min(kGJNP, total - r0);  // a multiple of 16
        if (r0 > 0) __syncthreads();  // the previous round's rows and sums are consumed
        // zero the eight staged rows of every slot (64 bytes each)
        for (int t = threadIdx.x; t < nslot * 8; t += NT) {
          const int s = t >> 3, r = t & 7;
          float4 *dst = r < 4 ? (float4 *)((r == 0 ? BH : r == 1 ? BL : r == 2 ? DH : DL) + s * kJRowB)
                              : (float4 *)((r == 4 ? WY : r == 5 ? DY : r == 6 ? WX : DX) +
                                           s * (r < 6 ? kJRowY : kJRowX));
#pragma unroll
          for (int i = 0; i < 4; ++i) dst[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        for (int s = threadIdx.x; s < nslot; s += NT) {
          const int li = LIST[r0 + s];
          const bool ok = li != 0xFFFF;
          const int gpos = r0 + (s & ~15);
          int c = 0;
#pragma unroll
          for (int k = 1; k < NCLS; ++k) c += gpos >= cstart[k];
          const float4 z = ok ? zr.r[2 * (base + li)] : make_float4(0.f, 0.f, 0.f, 0.f);
          const float4 rb = ok ? zr.r[2 * (base + li) + 1] : make_float4(0.f, 0.f, 0.f, 0.f);
          ZS[s] = z.x;
          ZS[kGJNP + s] = z.y;
          ZS[2 * kGJNP + s] = z.z;
          OFS[s] = ok ? __float_as_int(z.w) - (MW::origin(c) << 8) : 0;
          JIDX[s] = ok ? (int)__float_as_uint(rb.w) : -1;
          U[s] = make_float2(rb.x, rb.y);  // u travels in the records' value fields
        }
        __syncthreads();
        // taps: a task evaluates the value (d = 0) or derivative (d = 1) taps of
        // one axis for slots s, s + nslot / 2
        {
          const int hn = nslot >> 1;
          for (int u = threadIdx.x; u < 6 * hn; u += NT) {
            const int ad = u / hn, a = ad >> 1, d = ad & 1;
            const int s0 = u - ad * hn, s1 = s0 + hn;
            const int o0 = (OFS[s0] >> (8 * a)) & 255, o1 = (OFS[s1] >> (8 * a)) & 255;
            float v0[2 * W], v1[2 * W];
            if (d) taps2_sym<W, true>(PLD, ZS[a * kGJNP + s0], ZS[a * kGJNP + s1], v0, v1);
            else taps2_sym<W>(PL, ZS[a * kGJNP + s0], ZS[a * kGJNP + s1], v0, v1);
            if (a == 2) {
              store_split_rows<2 * W>(d ? DH : BH, d ? DL : BL, s0 * kJRowB + o0, s1 * kJRowB + o1,
                                      v0, v1, d ? kOpScaleD : kOpScale);
            } else {
              float *tab = a == 0 ? (d ? DX : WX) : (d ? DY : WY);
              const int stride = a == 0 ? kJRowX : kJRowY;
#pragma unroll
              for (int t = 0; t < 2 * W; ++t) {
                tab[s0 * stride + o0 + t] = v0[t];
                tab[s1 * stride + o1 + t] = v1[t];
              }
            }
          }
        }
        __syncthreads();
        for (int gi = 0; gi < (nslot >> 4); ++gi) {
          const int gpos = r0 + 16 * gi;
          int c = 0;
#pragma unroll
          for (int k = 1; k < NCLS; ++k) c += gpos >= cstart[k];
          if (c != wcls) {
            wcls = c;
            const int o = MW::origin(c);
            const int addr = 4 * (16 * g4 + ((i16 + o) & 15));
#pragma unroll
            for (int q = 0; q < XW; ++q) {
              wrh[q] = window_rows(trh[q][0], trh[q][1], o, i16, addr);
              wrl[q] = window_rows(trl[q][0], trl[q][1], o, i16, addr);
              wih[q] = window_rows(tih[q][0], tih[q][1], o, i16, addr);
              wil[q] = window_rows(til[q][0], til[q][1], o, i16, addr);
            }
          }
          const int k = 16 * gi + i16;  // this lane's slot
          const half8 bh = *(const half8 *)(BH + k * kJRowB + 8 * g4);
          const half8 bl = *(const half8 *)(BL + k * kJRowB + 8 * g4);
          const half8 dh = *(const half8 *)(DH + k * kJRowB + 8 * g4);
          const half8 dl = *(const half8 *)(DL + k * kJRowB + 8 * g4);
          // wy, wy' at this lane's D rows (window rows 4 g4 .. 4 g4 + 3); wx, wx'
          // at the wave's cells
          const float4 wy = *(const float4 *)(WY + k * kJRowY + 4 * g4);
          const float4 dy = *(const float4 *)(DY + k * kJRowY + 4 * g4);
          const float2 wx = *(const float2 *)(WX + k * kJRowX + wave * XW);
          const float2 dx = *(const float2 *)(DX + k * kJRowX + wave * XW);
          const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
          f32x4 dr[XW], di[XW], er[XW], ei[XW];
#pragma unroll
          for (int q = 0; q < XW; ++q) {
            dr[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wrh[q], bh, zero, 0, 0, 0);
            di[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wih[q], bh, zero, 0, 0, 0);
            er[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wrh[q], dh, zero, 0, 0, 0);
            ei[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wih[q], dh, zero, 0, 0, 0);
            dr[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wrh[q], bl, dr[q], 0, 0, 0);
            di[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wih[q], bl, di[q], 0, 0, 0);
            er[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wrh[q], dl, er[q], 0, 0, 0);
            ei[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wih[q], dl, ei[q], 0, 0, 0);
            dr[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wrl[q], bh, dr[q], 0, 0, 0);
            di[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wil[q], bh, di[q], 0, 0, 0);
            er[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wrl[q], dh, er[q], 0, 0, 0);
            ei[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wil[q], dh, ei[q], 0, 0, 0);
          }
          float sv[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};  // Sx, Sy, Sz (re, im)
#pragma unroll
          for (int q = 0; q < XW; ++q) {
            const float xw = q == 0 ? wx.x : wx.y, xd = q == 0 ? dx.x : dx.y;
            const f32x4 &r = dr[q], &i = di[q];
            const float ar = fmaf(wy.x, r[0], fmaf(wy.y, r[1], fmaf(wy.z, r[2], wy.w * r[3])));
            const float ai = fmaf(wy.x, i[0], fmaf(wy.y, i[1], fmaf(wy.z, i[2], wy.w * i[3])));
            const float br = fmaf(dy.x, r[0], fmaf(dy.y, r[1], fmaf(dy.z, r[2], dy.w * r[3])));
            const float bi = fmaf(dy.x, i[0], fmaf(dy.y, i[1], fmaf(dy.z, i[2], dy.w * i[3])));
            const float cr = fmaf(wy.x, er[q][0], fmaf(wy.y, er[q][1], fmaf(wy.z, er[q][2], wy.w * er[q][3])));
            const float ci = fmaf(wy.x, ei[q][0], fmaf(wy.y, ei[q][1], fmaf(wy.z, ei[q][2], wy.w * ei[q][3])));
            sv[0] = fmaf(xd, ar, sv[0]);
            sv[1] = fmaf(xd, ai, sv[1]);
            sv[2] = fmaf(xw, br, sv[2]);
            sv[3] = fmaf(xw, bi, sv[3]);
            sv[4] = fmaf(xw, cr, sv[4]);
            sv[5] = fmaf(xw, ci, sv[5]);
          }
          // lane rows g4 and g4 + 2 are added here, rows 0 / 1 in the output pass
#pragma unroll
          for (int c6 = 0; c6 < 6; ++c6) {
            sv[c6] += __shfl_xor(sv[c6], 32);
            if (g4 < 2) RED[((wave * 2 + g4) * 6 + c6) * kGJNP + k] = sv[c6];
          }
        }
        __syncthreads();
        for (int s = threadIdx.x; s < nslot; s += NT) {
          const int j = JIDX[s];
          if (j < 0) continue;
          float sv[6];
#pragma unroll
          for (int c6 = 0; c6 < 6; ++c6) {
            float a = 0.f;
#pragma unroll
            for (int r = 0; r < 2 * NW; ++r) a += RED[(r * 6 + c6) * kGJNP + s];
            sv[c6] = a * s_out * (c6 < 4 ? kProdUnscale : kProdUnscaleD);
          }
          const float2 uj = U[s];
          // Re(conj(u) * (-s) * S) = -s (u.x S.x + u.y S.y)
          gx[j] = -(float)g.scale[0] * (uj.x * sv[0] + uj.y * sv[1]);
          if (gy) gy[j] = -(float)g.scale[1] * (uj.x * sv[2] + uj.y * sv[3]);
          if (gz) gz[j] = -(float)g.scale[2] * (uj.x * sv[4] + uj.y * sv[5]);
        }
      }
    }
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
// ---- 1D: binned tiles of one wave --------------------------------------------
//
// Replaces the 1D loops of apollo-nufft (type1_into spread, transform/
// dimension_1d.rs:186-197; type2_into gather, :261-275) after the points are
// binned (sort_points_1d / sort_positions_1d, :54-99, replaced by the
// counting sort on 1D tiles).  The tile of bin b covers k1Tile cells from
// x0 = b bin - W + 1, with bin = k1Tile - 2W + 1, so every tap of its points
// lies inside.  A 64-thread workgroup takes subproblems of the list in turn
// (grid stride); lane j owns tile cell j.  Points come from the records of
// the binning (polynomial argument z, cell offset in the tile, value or u,
// input index).
constexpr int k1Tile = kTile1d;

// The 2W tap weights (and derivatives) at polynomial argument z, with
// kb_taps's endpoint rule.
template <int W>
__device__ __forceinline__ void taps_z(const KbPoly &P, float z, float (&wt)[2 * W]) {
#pragma unroll
  for (int i = 0; i < 2 * W; ++i) {
    float v = P.c[kPolyCoefs - 1][i];
#pragma unroll
    for (int d = kPolyCoefs - 2; d >= 0; --d) v = fmaf(v, z, P.c[d][i]);
    wt[i] = v;
  }
  if (z <= -1.0f) wt[0] = 0.0f;
  if (z >= 1.0f) wt[2 * W - 1] = 0.0f;
}

template <int W>
__device__ __forceinline__ void taps_z_deriv(const KbPoly &P, float z, float (&dwt)[2 * W]) {
#pragma unroll
  for (int i = 0; i < 2 * W; ++i) {
    float v = P.dc[kPolyCoefs - 1][i];
#pragma unroll
    for (int d = kPolyCoefs - 2; d >= 0; --d) v = fmaf(v, z, P.dc[d][i]);
    dwt[i] = v;
  }
  if (z <= -1.0f) dwt[0] = 0.0f;
  if (z >= 1.0f) dwt[2 * W - 1] = 0.0f;
}

// Spreading: the subproblem's points are staged 64 at a time (weights,
// values, offsets in LDS); each lane then adds, in list order, the
// contributions of the staged points whose taps cover its cell, and the
// finished tile is added to the grid (cells shared with neighbouring tiles
// by atomics).
template <int W>
__global__ __launch_bounds__(k1Tile) void spread1d_binned_kernel(PointArgs pa, ZRec zr,
                                                                 float2 *grid, Geom g,
                                                                 KbPoly poly) {
  constexpr int T = 2 * W;
  __shared__ float wts[k1Tile][T + 1];
  __shared__ float2 cv[k1Tile];
  __shared__ int ofs[k1Tile];
  const int lane = threadIdx.x, m = g.m[0];
  const uint32_t nsub = pa.nsub[0];
  for (uint32_t si = blockIdx.x; si < nsub; si += gridDim.x) {
    const int4 sub = pa.subs[si];
    const int x0 = sub.x * g.bin[0] - W + 1;
    float ar = 0.f, ai = 0.f;
    for (int c0 = 0; c0 < sub.z; c0 += k1Tile) {
      const int np = min(k1Tile, sub.z - c0);
      __syncthreads();  // the previous batch is consumed
      if (lane < np) {
        const int64_t r = (int64_t)sub.y + c0 + lane;
        const float4 ra = zr.r[2 * r], rb = zr.r[2 * r + 1];
        float wt[T];
        taps_z<W>(poly, ra.x, wt);
#pragma unroll
        for (int i = 0; i < T; ++i) wts[lane][i] = wt[i];
        cv[lane] = make_float2(rb.x, rb.y);
        ofs[lane] = __float_as_int(ra.w) & 255;
      }
      __syncthreads();
      for (int q = 0; q < np; ++q) {
        const int d = lane - ofs[q];
        if (d >= 0 && d < T) {
          const float wv = wts[q][d];
          const float2 c = cv[q];
          ar = fmaf(c.x, wv, ar);
          ai = fmaf(c.y, wv, ai);
        }
      }
    }
    if (ar != 0.f || ai != 0.f) {
      const int x = wrap(x0 + lane, m);
      unsafeAtomicAdd(&grid[x].x, ar);
      unsafeAtomicAdd(&grid[x].y, ai);
    }
  }
}

// Interpolation (DERIV: the coordinate gradient): the subproblem's tile of
// the grid in LDS, then lane p takes point p of each batch of 64.
template <int W, bool DERIV>
__global__ __launch_bounds__(k1Tile) void interp1d_binned_kernel(PointArgs pa, ZRec zr,
                                                                 const float2 *grid, float2 *out,
                                                                 float *gx, Geom g,
                                                                 KbPoly poly) {
  constexpr int T = 2 * W;
  __shared__ float2 tile[k1Tile];
  const int lane = threadIdx.x, m = g.m[0];
  const uint32_t nsub = pa.nsub[0];
  for (uint32_t si = blockIdx.x; si < nsub; si += gridDim.x) {
    const int4 sub = pa.subs[si];
    const int x0 = sub.x * g.bin[0] - W + 1;
    __syncthreads();  // the previous tile is consumed
    tile[lane] = grid[wrap(x0 + lane, m)];
    __syncthreads();
    for (int c0 = lane; c0 < sub.z; c0 += k1Tile) {
      const int64_t r = (int64_t)sub.y + c0;
      const float4 ra = zr.r[2 * r], rb = zr.r[2 * r + 1];
      const int o = __float_as_int(ra.w) & 255;
      float wt[T];
      if (DERIV) taps_z_deriv<W>(poly, ra.x, wt);
      else taps_z<W>(poly, ra.x, wt);
      float sr = 0.f, si2 = 0.f;
#pragma unroll
      for (int i = 0; i < T; ++i) {
        const float2 v = tile[o + i];
        sr = fmaf(wt[i], v.x, sr);
        si2 = fmaf(wt[i], v.y, si2);
      }
      const uint32_t j = __float_as_uint(rb.w);
      if (DERIV) gx[j] = -(float)g.scale[0] * (rb.x * sr + rb.y * si2);
      else out[j] = make_float2(sr, si2);
    }
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
struct PointPrep {
  static hipError_t run(const PointArgs &pa, const float2 *values, const Geom &g,
                        const KbPoly &poly, float *rec, uint8_t *xo, hipStream_t s) {
    if (pa.num_points <= 0) return hipSuccess;
    int64_t blocks = (pa.num_points + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(point_prep_kernel<W>, dim3((unsigned)blocks), dim3(256), 0, s, pa, values,
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
struct Spread3Mfma {
  static hipError_t run(const PointArgs &pa, ZRec zr, float2 *grid, const Geom &g,
                        const KbPoly *dpoly, int nblk, hipStream_t s) {
    constexpr int lds = MLds<MSpread<W>::PX>::SIZE;
    hipError_t e = hipFuncSetAttribute((const void *)spread3d_mfma_kernel<W>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(spread3d_mfma_kernel<W>, dim3(nblk), dim3(MSpread<W>::NT), lds, s, pa, zr,
                       grid, g, dpoly);
    return hipGetLastError();
  }
};

template <int W>
struct Spread3MfmaBatched {
  static hipError_t run(const SpreadProblem *probs, const uint32_t *prefix, int nprob,
                        const KbPoly *poly, int nblk, hipStream_t s) {
    constexpr int lds = MLds<MSpread<W>::PX>::SIZE;
    hipError_t e = hipFuncSetAttribute((const void *)spread3d_mfma_batched_kernel<W>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(spread3d_mfma_batched_kernel<W>, dim3(nblk), dim3(MSpread<W>::NT), lds,
                       s, probs, prefix, nprob, poly);
    return hipGetLastError();
  }
};

template <int W>
struct SpreadBinX {
  static hipError_t run(int method, int *out) {
    *out = method == kSpreadMfma ? MSpread<W>::BINX : kSpreadBinX;
    return hipSuccess;
  }
};

template <int W>
struct MSpreadLds {
  static hipError_t run(size_t *out) {
    *out = std::max<size_t>(MLds<MSpread<W>::PX>::SIZE, JLds<MSpread<W>::PX, MSpread<W>::NW>::SIZE);
    return hipSuccess;
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
struct TileMax3 {
  static hipError_t run(const float2 *grid, float *bmax, const Geom &g, hipStream_t s,
                        bool bin_max_done) {
    // bin maxima (first half of the scratch; unless the last FFT pass wrote
    // them) -> tile maxima (second half)
    const int nbins = g.nbin[0] * g.nbin[1] * g.nbin[2];
    float *tmax = bmax + nbins;
    // tile_max_kernel lists at most kTileMaxList bins per axis
    if (MSpread<W>::PX / g.bin[0] + 2 > kTileMaxList || g.pad[1] / g.bin[1] + 2 > kTileMaxList ||
        g.pad[2] / g.bin[2] + 2 > kTileMaxList)
      return hipErrorInvalidValue;
    if (!bin_max_done)
      hipLaunchKernelGGL(bin_max_kernel, dim3(g.nbin[0] * g.nbin[1]), dim3(256),
                         sizeof(uint32_t) * g.nbin[2], s, grid, g, bmax);
    hipLaunchKernelGGL(tile_max_kernel, dim3(nbins), dim3(64), 0, s, (const float *)bmax, g,
                       make_int3(1 - W, -W, -W), make_int3(MSpread<W>::PX, g.pad[1], g.pad[2]), tmax);
    return hipGetLastError();
  }
};

template <int W>
struct Interp3Mfma {
  static hipError_t run(const PointArgs &pa, ZRec zr, const float2 *grid, float *bmax,
                        float2 *out, const Geom &g, const KbPoly *dpoly, int nblk, int cus,
                        hipStream_t s) {
    const float *tmax = bmax + g.nbin[0] * g.nbin[1] * g.nbin[2];  // from TileMax3
    // older / younger shares only when the grid holds two workgroups per CU
    const bool two = win_groups_per_cu<W>() == 2 && nblk == 2 * cus;
    const uint32_t wa = two ? kInterpOlder : 1u, wb = two ? kInterpYounger : 1u;
    // 16-row windows hold the 2w tap rows of a class of points
    constexpr int lds = JLds<MSpread<W>::PX, MSpread<W>::NW>::SIZE;
    hipError_t e = hipFuncSetAttribute((const void *)interp3d_win_kernel<W>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(interp3d_win_kernel<W>, dim3(nblk), dim3(MSpread<W>::NT), lds, s, pa, zr,
                       grid, (const float *)tmax, out, g, dpoly, wa, wb);
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
struct Grad3Mfma {
  static hipError_t run(const PointArgs &pa, ZRec zr, const float2 *grid, const float2 *u,
                        float *gx, float *gy, float *gz, const Geom &g, const KbPoly *dpoly,
                        int nblk, hipStream_t s) {
    if constexpr (W < 8) {
      constexpr int lds = GJLds<MGrad<W>::NW>::SIZE;
      hipError_t e = hipFuncSetAttribute((const void *)grad3d_win_kernel<W>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, lds);
      if (e != hipSuccess) return e;
      // u reaches the kernel through the records' value fields
      (void)u;
      hipLaunchKernelGGL(grad3d_win_kernel<W>, dim3(nblk), dim3(MGrad<W>::NT), lds, s, pa, zr,
                         grid, gx, gy, gz, g, dpoly);
      return hipGetLastError();
    } else {
      return hipErrorInvalidValue;  // w = 8 uses the LDS-tile kernel (grad_uses_mfma)
    }
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

template <int W>
struct Spread1 {
  static hipError_t run(const PointArgs &pa, ZRec zr, float2 *grid, const Geom &g,
                        const KbPoly &poly, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(spread1d_binned_kernel<W>, dim3(blocks), dim3(k1Tile), 0, s, pa, zr, grid,
                       g, poly);
    return hipGetLastError();
  }
};

template <int W>
struct Interp1 {
  static hipError_t run(const PointArgs &pa, ZRec zr, const float2 *grid, float2 *out,
                        float *gx, const Geom &g, const KbPoly &poly, int blocks, hipStream_t s) {
    if (gx)
      hipLaunchKernelGGL((interp1d_binned_kernel<W, true>), dim3(blocks), dim3(k1Tile), 0, s, pa,
                         zr, grid, out, gx, g, poly);
    else
      hipLaunchKernelGGL((interp1d_binned_kernel<W, false>), dim3(blocks), dim3(k1Tile), 0, s, pa,
                         zr, grid, out, gx, g, poly);
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

int point_rec_floats(int w) { return (2 + 6 * w + 1 + 3) / 4 * 4; }


hipError_t launch_point_prep(int w, const PointArgs &pa, const float2 *values, const Geom &g,
                             const KbPoly &poly, float *rec, uint8_t *xo, hipStream_t s) {
  return dispatch_w<PointPrep>(w, pa, values, g, poly, rec, xo, s);
}

int spread_bin_x(int w, int method) {
  int v = kSpreadBinX;
  (void)dispatch_w<SpreadBinX>(w, method, &v);
  return v;
}

int interp_blocks_per_cu(int w, int method) {
  if (method != kSpreadMfma) return 4;
  return w <= 7 ? 2 : 1;  // win_groups_per_cu
}

int spread_block_threads(int w, int method) {
  if (method != kSpreadMfma) return kSpreadThreads;
  return w >= 8 ? 512 : 256;
}

hipError_t launch_spread3d(int w, int method, const PointArgs &pa, const float *rec,
                           const uint8_t *xo, ZRec zr, float2 *grid, const Geom &g,
                           const KbPoly &poly, const KbPoly *dpoly, int nblk, hipStream_t s) {
  if (method == kSpreadMfma) return dispatch_w<Spread3Mfma>(w, pa, zr, grid, g, dpoly, nblk, s);
  return dispatch_w<Spread3>(w, pa, rec, xo, grid, g, poly, nblk, s);
}
hipError_t launch_spread3d_batched(int w, int method, const SpreadProblem *host_probs, int nprob,
                                   SpreadProblem *dev_table, uint32_t *dev_prefix,
                                   const KbPoly &poly, const KbPoly *dpoly, int nblk,
                                   hipStream_t s) {
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
  if (method == kSpreadMfma)
    return dispatch_w<Spread3MfmaBatched>(w, (const SpreadProblem *)dev_table,
                                          (const uint32_t *)dev_prefix, nprob, dpoly, nblk, s);
  return dispatch_w<Spread3Batched>(w, (const SpreadProblem *)dev_table, (const uint32_t *)dev_prefix,
                                    nprob, poly, nblk, s);
}

hipError_t launch_tile_max3d(int w, int method, const float2 *grid, float *scratch, const Geom &g,
                             hipStream_t s, bool bin_max_done) {
  if (method != kSpreadMfma) return hipSuccess;
  return dispatch_w<TileMax3>(w, grid, scratch, g, s, bin_max_done);
}
hipError_t launch_interp3d(int w, int method, const PointArgs &pa, ZRec zr, const float2 *grid,
                           float *scratch, float2 *out, const Geom &g, const KbPoly &poly,
                           const KbPoly *dpoly, int nblk, int cus, hipStream_t s) {
  if (method == kSpreadMfma)
    return dispatch_w<Interp3Mfma>(w, pa, zr, grid, scratch, out, g, dpoly, nblk, cus, s);
  return dispatch_w<Interp3>(w, pa, grid, out, g, poly, nblk, s);
}
hipError_t launch_grad3d(int w, int method, const PointArgs &pa, ZRec zr, const float2 *grid,
                         const float2 *u, float *gx, float *gy, float *gz, const Geom &g,
                         const KbPoly &poly, const KbPoly *dpoly, int nblk, hipStream_t s) {
  if (grad_uses_mfma(w, method))
    return dispatch_w<Grad3Mfma>(w, pa, zr, grid, u, gx, gy, gz, g, dpoly, nblk, s);
  return dispatch_w<Grad3>(w, pa, grid, u, gx, gy, gz, g, poly, nblk, s);
}
hipError_t launch_spread1d(int w, const PointArgs &pa, ZRec zr, float2 *grid, const Geom &g,
                           const KbPoly &poly, int num_blocks, hipStream_t s) {
  return dispatch_w<Spread1>(w, pa, zr, grid, g, poly, num_blocks, s);
}
hipError_t launch_interp1d(int w, const PointArgs &pa, ZRec zr, const float2 *grid, float2 *out,
                           float *gx, const Geom &g, const KbPoly &poly, int num_blocks,
                           hipStream_t s) {
  return dispatch_w<Interp1>(w, pa, zr, grid, out, gx, g, poly, num_blocks, s);
}
size_t spread_lds_bytes(int w, int method, const Geom &g) {
  if (method == kSpreadMfma) {
    size_t v = 0;
    (void)dispatch_w<MSpreadLds>(w, &v);
    return v;
  }
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
