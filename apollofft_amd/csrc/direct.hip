// direct.hip -- exact direct NUFFT sums on the device in FP64.
//
// The reference's WGPU "direct" backend (apollo-nufft-wgpu
// infrastructure/shaders/nufft.wgsl:51-140, driven by kernel.rs:652-812)
// evaluates
//   Type-1: F_k = sum_j c_j exp(-2 pi i sum_a k_a x_{j,a} / L_a)   (every mode)
//   Type-2: g_j = sum_k f_k exp(+2 pi i sum_a k_a x_{j,a} / L_a)   (every point)
// with k_a the signed FFT-order index (k <= n/2 -> k, else k - n), in f32.
// Here the same sums run in FP64 (float32 coordinates and complex64 data
// are widened once; results are complex128), which also makes them a fast
// device oracle for large inputs.  The host-pointer entry points (the exact
// functions of the Python binding, apollo-python lib.rs:754-836) pass FP64
// coordinates and complex128 data through unchanged.
//
// Type-1: one thread per mode; points are staged 256 at a time in LDS as
// their three FP64 phases theta_a = -2 pi x_a / L_a and value, and each
// thread adds c_j (cos a, sin a) with a = sum_a k_a theta_a, one FP64
// sincos per (mode, point).
// Type-2: one thread per point; modes are walked line by line along the
// last axis, each line split into its non-negative and negative index runs;
// the factor exp(i k theta) of a run starts from one exact sincos and
// advances by complex multiplication (FP64; the drift over a run of n/2
// steps stays near n/2 ulp).  The mode lines are staged in LDS and read by
// every thread of the block (broadcast).
#include "internal.h"

namespace anufft {
namespace {

constexpr int kDirThreads = 256;
constexpr int kDirLine = 1024;  // mode-line chunk staged per round (complex)

__device__ __forceinline__ int64_t signed_idx(int64_t k, int64_t n) {
  return k <= n / 2 ? k : k - n;
}

template <typename TC>
struct DirectArgs {
  const TC *x, *y, *z;
  int64_t stride;
  int64_t num_points;
  int dim;
  int64_t n[3];
  double twopi_over_l[3];  // 2 pi / L_a
};

template <typename TC, typename TV>
__global__ __launch_bounds__(kDirThreads) void direct_type1_kernel(DirectArgs<TC> a,
                                                                   const TV *values,
                                                                   double2 *out) {
  __shared__ double th[3][kDirThreads];
  __shared__ double2 cv[kDirThreads];
  const int64_t N = a.n[0] * a.n[1] * a.n[2];
  const int64_t f = (int64_t)blockIdx.x * kDirThreads + threadIdx.x;
  double k0 = 0.0, k1 = 0.0, k2 = 0.0;
  if (f < N) {
    const int64_t i2 = f % a.n[2], i1 = (f / a.n[2]) % a.n[1], i0 = f / (a.n[1] * a.n[2]);
    k0 = (double)signed_idx(i0, a.n[0]);
    k1 = (double)signed_idx(i1, a.n[1]);
    k2 = (double)signed_idx(i2, a.n[2]);
  }
  double sr = 0.0, si = 0.0;
  for (int64_t j0 = 0; j0 < a.num_points; j0 += kDirThreads) {
    const int64_t j = j0 + threadIdx.x;
    __syncthreads();  // the previous tile is consumed
    if (j < a.num_points) {
      th[0][threadIdx.x] = -a.twopi_over_l[0] * (double)a.x[j * a.stride];
      th[1][threadIdx.x] = a.dim == 3 ? -a.twopi_over_l[1] * (double)a.y[j * a.stride] : 0.0;
      th[2][threadIdx.x] = a.dim == 3 ? -a.twopi_over_l[2] * (double)a.z[j * a.stride] : 0.0;
      const TV c = values[j];
      cv[threadIdx.x] = make_double2((double)c.x, (double)c.y);
    }
    __syncthreads();
    const int nt = (int)min((int64_t)kDirThreads, a.num_points - j0);
    if (f < N) {
      for (int t = 0; t < nt; ++t) {
        const double ang = fma(k0, th[0][t], fma(k1, th[1][t], k2 * th[2][t]));
        double s, c;
        sincos(ang, &s, &c);
        const double2 v = cv[t];
        sr = fma(v.x, c, fma(-v.y, s, sr));
        si = fma(v.x, s, fma(v.y, c, si));
      }
    }
  }
  if (f < N) out[f] = make_double2(sr, si);
}

template <typename TC, typename TV>
__global__ __launch_bounds__(kDirThreads) void direct_type2_kernel(DirectArgs<TC> a,
                                                                   const TV *modes,
                                                                   double2 *out) {
  __shared__ double2 line[kDirLine];
  const int64_t j = (int64_t)blockIdx.x * kDirThreads + threadIdx.x;
  const bool ok = j < a.num_points;
  double th0 = 0.0, th1 = 0.0, th2 = 0.0;
  if (ok) {
    th0 = a.twopi_over_l[0] * (double)a.x[j * a.stride];
    if (a.dim == 3) {
      th1 = a.twopi_over_l[1] * (double)a.y[j * a.stride];
      th2 = a.twopi_over_l[2] * (double)a.z[j * a.stride];
    }
  }
  // the last axis carries the runs: for 1D plans that is axis 0
  const int64_t nl = a.dim == 3 ? a.n[2] : a.n[0];
  const double thl = a.dim == 3 ? th2 : th0;
  const int64_t nouter = a.dim == 3 ? a.n[0] * a.n[1] : 1;
  double sr = 0.0, si = 0.0;
  for (int64_t o = 0; o < nouter; ++o) {
    // phase of the line: k0 th0 + k1 th1 (3D)
    double base = 0.0;
    if (a.dim == 3) {
      const int64_t i0 = o / a.n[1], i1 = o - i0 * a.n[1];
      base = (double)signed_idx(i0, a.n[0]) * th0 + (double)signed_idx(i1, a.n[1]) * th1;
    }
    double bs, bc;
    sincos(base, &bs, &bc);
    double lr = 0.0, li = 0.0;  // sum over the line of f_k exp(i k thl)
    for (int64_t c0 = 0; c0 < nl; c0 += kDirLine) {
      const int cl = (int)min((int64_t)kDirLine, nl - c0);
      __syncthreads();  // the previous chunk is consumed
      for (int t = threadIdx.x; t < cl; t += kDirThreads) {
        const TV v = modes[o * nl + c0 + t];
        line[t] = make_double2((double)v.x, (double)v.y);
      }
      __syncthreads();
      if (ok) {
        // runs of consecutive signed indices inside the chunk: [c0, min(end, n/2]]
        // with k = index, and [max(c0, n/2 + 1), end) with k = index - n
        const int64_t split = nl / 2 + 1;  // first index with a negative k
        for (int run = 0; run < 2; ++run) {
          const int64_t r0 = run == 0 ? c0 : max(c0, split);
          const int64_t r1 = run == 0 ? min(c0 + (int64_t)cl, split) : c0 + (int64_t)cl;
          if (r0 >= r1) continue;
          double es, ec;  // exp(i k thl) at the run start
          sincos((double)signed_idx(r0, nl) * thl, &es, &ec);
          double ss, sc;  // one step
          sincos(thl, &ss, &sc);
          for (int64_t i = r0; i < r1; ++i) {
            const double2 v = line[i - c0];
            lr = fma(v.x, ec, fma(-v.y, es, lr));
            li = fma(v.x, es, fma(v.y, ec, li));
            const double nc = ec * sc - es * ss, ns = fma(ec, ss, es * sc);
            ec = nc;
            es = ns;
          }
        }
      }
    }
    // times exp(i base)
    sr = fma(lr, bc, fma(-li, bs, sr));
    si = fma(lr, bs, fma(li, bc, si));
  }
  if (ok) out[j] = make_double2(sr, si);
}

// Type-1 at selected modes (flat FFT-order indices): blockIdx.x covers 256
// modes, blockIdx.y one run of points; each block adds its partial sums to
// out with FP64 atomics (out zeroed by zero_double2_kernel first).
constexpr int kSelChunk = 65536;  // points per block row

__global__ __launch_bounds__(kDirThreads) void zero_double2_kernel(double2 *p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kDirThreads + threadIdx.x;
  if (i < n) p[i] = make_double2(0.0, 0.0);
}

__global__ __launch_bounds__(kDirThreads) void direct_type1_select_kernel(
    DirectArgs<float> a, const double2 *values, const int64_t *sel, int64_t nsel, double2 *out) {
  __shared__ double th[3][kDirThreads];
  __shared__ double2 cv[kDirThreads];
  const int64_t s = (int64_t)blockIdx.x * kDirThreads + threadIdx.x;
  double k0 = 0.0, k1 = 0.0, k2 = 0.0;
  if (s < nsel) {
    const int64_t f = sel[s];
    const int64_t i2 = f % a.n[2], i1 = (f / a.n[2]) % a.n[1], i0 = f / (a.n[1] * a.n[2]);
    k0 = (double)signed_idx(i0, a.n[0]);
    k1 = (double)signed_idx(i1, a.n[1]);
    k2 = (double)signed_idx(i2, a.n[2]);
  }
  const int64_t jlo = (int64_t)blockIdx.y * kSelChunk;
  const int64_t jhi = min(a.num_points, jlo + kSelChunk);
  double sr = 0.0, si = 0.0;
  for (int64_t j0 = jlo; j0 < jhi; j0 += kDirThreads) {
    const int64_t j = j0 + threadIdx.x;
    __syncthreads();
    if (j < jhi) {
      th[0][threadIdx.x] = -a.twopi_over_l[0] * (double)a.x[j * a.stride];
      th[1][threadIdx.x] = a.dim == 3 ? -a.twopi_over_l[1] * (double)a.y[j * a.stride] : 0.0;
      th[2][threadIdx.x] = a.dim == 3 ? -a.twopi_over_l[2] * (double)a.z[j * a.stride] : 0.0;
      cv[threadIdx.x] = values[j];
    }
    __syncthreads();
    const int nt = (int)min((int64_t)kDirThreads, jhi - j0);
    if (s < nsel) {
      for (int t = 0; t < nt; ++t) {
        const double ang = fma(k0, th[0][t], fma(k1, th[1][t], k2 * th[2][t]));
        double sn, cs;
        sincos(ang, &sn, &cs);
        const double2 v = cv[t];
        sr = fma(v.x, cs, fma(-v.y, sn, sr));
        si = fma(v.x, sn, fma(v.y, cs, si));
      }
    }
  }
  if (s < nsel) {
    unsafeAtomicAdd(&out[s].x, sr);
    unsafeAtomicAdd(&out[s].y, si);
  }
}

template <typename TC, typename TV>
hipError_t launch_direct_t(int type, const TC *x, const TC *y, const TC *z, int64_t stride,
                           int64_t num_points, int dim, const int64_t n[3], const double L[3],
                           const TV *in, double2 *out, hipStream_t s) {
  DirectArgs<TC> a{};
  a.x = x;
  a.y = y;
  a.z = z;
  a.stride = stride;
  a.num_points = num_points;
  a.dim = dim;
  for (int i = 0; i < 3; ++i) {
    a.n[i] = dim == 3 || i == 0 ? n[i] : 1;
    a.twopi_over_l[i] = 2.0 * M_PI / L[i];
  }
  if (type == 1) {
    const int64_t N = a.n[0] * a.n[1] * a.n[2];
    hipLaunchKernelGGL((direct_type1_kernel<TC, TV>),
                       dim3((unsigned)((N + kDirThreads - 1) / kDirThreads)), dim3(kDirThreads), 0,
                       s, a, in, out);
  } else {
    if (num_points <= 0) return hipSuccess;
    hipLaunchKernelGGL((direct_type2_kernel<TC, TV>),
                       dim3((unsigned)((num_points + kDirThreads - 1) / kDirThreads)),
                       dim3(kDirThreads), 0, s, a, in, out);
  }
  return hipGetLastError();
}

} // namespace

hipError_t launch_direct(int type, const PointArgs &pa, int dim, const int64_t n[3],
                         const double L[3], const float2 *in, double2 *out, hipStream_t s) {
  return launch_direct_t<float, float2>(type, pa.x, pa.y, pa.z, pa.stride, pa.num_points, dim, n,
                                        L, in, out, s);
}

hipError_t launch_direct_f64(int type, const double *x, const double *y, const double *z,
                             int64_t stride, int64_t num_points, int dim, const int64_t n[3],
                             const double L[3], const double2 *in, double2 *out, hipStream_t s) {
  return launch_direct_t<double, double2>(type, x, y, z, stride, num_points, dim, n, L, in, out,
                                          s);
}

hipError_t launch_direct_select(const PointArgs &pa, int dim, const int64_t n[3],
                                const double L[3], const double2 *values, const int64_t *sel,
                                int64_t nsel, double2 *out, hipStream_t s) {
  if (nsel <= 0) return hipSuccess;
  hipLaunchKernelGGL(zero_double2_kernel, dim3((unsigned)((nsel + kDirThreads - 1) / kDirThreads)),
                     dim3(kDirThreads), 0, s, out, nsel);
  if (pa.num_points <= 0) return hipGetLastError();
  DirectArgs<float> a{};
  a.x = pa.x;
  a.y = pa.y;
  a.z = pa.z;
  a.stride = pa.stride;
  a.num_points = pa.num_points;
  a.dim = dim;
  for (int i = 0; i < 3; ++i) {
    a.n[i] = dim == 3 || i == 0 ? n[i] : 1;
    a.twopi_over_l[i] = 2.0 * M_PI / L[i];
  }
  const dim3 grid((unsigned)((nsel + kDirThreads - 1) / kDirThreads),
                  (unsigned)((pa.num_points + kSelChunk - 1) / kSelChunk));
  hipLaunchKernelGGL(direct_type1_select_kernel, grid, dim3(kDirThreads), 0, s, a, values, sel,
                     nsel, out);
  return hipGetLastError();
}

} // namespace anufft
