// spread.hip -- Kaiser-Bessel spreading (Type-1), interpolation (Type-2) and
// the point-coordinate gradient, 3D subproblems staged in LDS plus 1D paths.
//
// Replaces the per-point T^3 loops of apollo-nufft (type1_into spread,
// transform/dimension_3d.rs:290-327; type2_into gather, :560-601; 1D
// dimension_1d.rs:186-197, :261-275) and the O(grid x points) WGSL passes
// fast_type1_spread_3d / fast_type2_interpolate_3d (apollo-nufft-wgpu
// .../shaders/nufft_fast_3d.wgsl:130-161, :268-299).
//
// 3D work unit = subproblem: up to max_sub points of one tile (bin[0] x
// bin[1] x bin[2] fine cells).  A block
//   1. stages 128 points at a time: coordinates -> grid units in FP64, the
//      2w tap weights of every axis from the per-tap polynomials (FP32),
//      values, and the local offset of the first tap inside the padded tile;
//   2. spread: each wave takes 16 points at once and maps its lanes to
//      (point, y-tap, z-tap) triples (16 (2w)^2 = 64 w^2 items, so every
//      lane is busy for every even tap count), loops over the 2w x-taps
//      and adds into the padded tile with LDS FP64 atomics (split re/im
//      planes, consecutive z-taps on consecutive banks).  On gfx950 the
//      LDS FP64 add runs ~23x faster than the LDS FP32 add (measured
//      7.8 vs 0.33 lane-ops/clk/CU), and the tile accumulates in FP64;
//   3. flush: the padded tile (bin + 2w per axis) is rounded to FP32 and
//      added to the HBM grid with global float atomics, skipping zero
//      cells, periodic wrap.
// Interpolation loads the padded tile from HBM once, each wave gathers one
// point at a time (lanes over (y,z) taps, loop over x) and reduces across
// the wave.
#include "internal.h"

namespace anufft {
namespace {

constexpr int kStage = 128;       // points staged per round (interp, grad)
constexpr int kSpreadStage = 64;  // points staged per round (spread)
constexpr int kSpreadThreads = 512;
constexpr int kInterpThreads = 256;

// Periodic index; padded tiles may exceed the grid by more than one period
// when the grid is not much larger than the kernel support.
__device__ __forceinline__ int wrap(int v, int m) {
  while (v < 0) v += m;
  while (v >= m) v -= m;
  return v;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int W>
struct Staging {
  static constexpr int T = 2 * W;
  static constexpr int WS = 3 * T + 1; // weights per point (+1 pad)
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
      i0 = kb_taps_deriv<W>(t, poly, wt, dwt);
#pragma unroll
      for (int i = 0; i < T; ++i) dwts[p * WS + ax * T + i] = dwt[i];
    } else {
      i0 = kb_taps<W>(t, poly, wt);
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

template <int W>
__global__ __launch_bounds__(kSpreadThreads) void spread3d_kernel(PointArgs pa,
                                                                  const float2 *values,
                                                                  float2 *grid, Geom g,
                                                                  KbPoly poly) {
  constexpr int T = 2 * W;
  constexpr int WS = Staging<W>::WS;
  constexpr int NT = kSpreadThreads;
  extern __shared__ double shd[];
  const int Px = g.pad[0], Py = g.pad[1], Pz = g.pad[2];
  const int P = Px * Py * Pz;
  const int PyPz = Py * Pz;
  double *tre = shd;
  double *tim = shd + P;
  float *wts = (float *)(shd + 2 * P);
  float2 *cv = (float2 *)(wts + ((kSpreadStage * WS + 1) & ~1));
  int *loc = (int *)(cv + kSpreadStage);

  for (int i = threadIdx.x; i < 2 * P; i += NT) shd[i] = 0.0;
  __syncthreads();

  const uint32_t nsub = *pa.nsub;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int kWaves = NT / 64;
  for (uint32_t sidx = blockIdx.x; sidx < nsub; sidx += gridDim.x) {
    const int4 sub = pa.subs[sidx];
    const int bin = sub.x;
    const int bz = bin % g.nbin[2];
    const int by = (bin / g.nbin[2]) % g.nbin[1];
    const int bx = bin / (g.nbin[2] * g.nbin[1]);
    const int ox = bx * g.bin[0] - W, oy = by * g.bin[1] - W, oz = bz * g.bin[2] - W;

    for (int c0 = 0; c0 < sub.z; c0 += kSpreadStage) {
      const int npts = min(kSpreadStage, sub.z - c0);
      stage_points<W, NT, false>(pa, g, poly, values, sub.y + c0, npts, ox, oy, oz, wts,
                                 nullptr, cv, loc, nullptr);
      __syncthreads();
      for (int g0 = wave * 16; g0 < npts; g0 += kWaves * 16) {
        for (int r = 0; r < W * W; ++r) {
          const int q = lane + 64 * r;
          const int pl = q / (T * T);
          const int rem = q - pl * (T * T);
          const int ty = rem / T, tz = rem - (rem / T) * T;
          const int p = g0 + pl;
          if (p < npts) {
            const float *wp = wts + p * WS;
            const float wyz = wp[T + ty] * wp[2 * T + tz];
            const float2 c = cv[p];
            const float cr = c.x * wyz, ci = c.y * wyz;
            const int base = (loc[3 * p] * Py + loc[3 * p + 1] + ty) * Pz + loc[3 * p + 2] + tz;
#pragma unroll
            for (int tx = 0; tx < T; ++tx) {
              const float wx = wp[tx];
              atomicAdd(&tre[base + tx * PyPz], (double)(cr * wx));
              atomicAdd(&tim[base + tx * PyPz], (double)(ci * wx));
            }
          }
        }
      }
      __syncthreads();
    }

    // flush the padded tile into the grid and clear it
    for (int i = threadIdx.x; i < P; i += NT) {
      const double re = tre[i], im = tim[i];
      if (re != 0.0 || im != 0.0) {
        tre[i] = 0.0;
        tim[i] = 0.0;
        const int pz = i % Pz;
        const int t2 = i / Pz;
        const int py = t2 % Py;
        const int px = t2 / Py;
        const int gx = wrap(ox + px, g.m[0]);
        const int gy = wrap(oy + py, g.m[1]);
        const int gz = wrap(oz + pz, g.m[2]);
        float2 *dst = grid + ((int64_t)gx * g.m[1] + gy) * g.m[2] + gz;
        unsafeAtomicAdd(&dst->x, (float)re);
        unsafeAtomicAdd(&dst->y, (float)im);
      }
    }
    __syncthreads();
  }
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
    int i0 = kb_taps<W>(t, poly, wt);
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
    int i0 = kb_taps<W>(t, poly, wt);
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
    int i0 = kb_taps_deriv<W>(t, poly, wt, dwt);
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
size_t spread_lds(const Geom &g) {
  const size_t P = (size_t)g.pad[0] * g.pad[1] * g.pad[2];
  const size_t ws = ((size_t)kSpreadStage * Staging<W>::WS + 1) & ~(size_t)1;
  return 2 * P * sizeof(double) + ws * sizeof(float) + kSpreadStage * sizeof(float2) +
         3 * kSpreadStage * sizeof(int);
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
struct Spread3 {
  static hipError_t run(const PointArgs &pa, const float2 *values, float2 *grid, const Geom &g,
                        const KbPoly &poly, int nblk, hipStream_t s) {
    size_t lds = spread_lds<W>(g);
    hipError_t e = hipFuncSetAttribute((const void *)spread3d_kernel<W>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(spread3d_kernel<W>, dim3(nblk), dim3(kSpreadThreads), lds, s, pa, values,
                       grid, g, poly);
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

hipError_t launch_spread3d(int w, const PointArgs &pa, const float2 *values, float2 *grid,
                           const Geom &g, const KbPoly &poly, int nblk, hipStream_t s) {
  return dispatch_w<Spread3>(w, pa, values, grid, g, poly, nblk, s);
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
