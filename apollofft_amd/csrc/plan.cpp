// plan.cpp -- plans, KB metadata, workspace layout and the C ABI of
// libapollo_nufft_hip (declared in include/apollo_nufft.h).
//
// Plan creation follows NufftPlan1D::new / NufftPlan3D::new
// (apollo-nufft/src/application/execution/transform/dimension_1d.rs:127-154,
// dimension_3d.rs:193-232): same oversampled length rule, same beta with the
// integer sigma, same FP64 deconvolution factors.  In addition it builds the
// device tables the kernels need: FP64 deconvolution factors, FP64-built
// twiddles rounded to FP32, and per-tap polynomial fits of the KB kernel and
// its derivative.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/apollo_nufft.h"
#include "internal.h"

using namespace anufft;

namespace {

thread_local std::string g_last_error;

anufft_status fail(anufft_status st, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return st;
}

anufft_status hip_fail(hipError_t e, const char *what) {
  return fail(ANUFFT_HIP_ERROR, "%s: %s", what, hipGetErrorString(e));
}

// ---- KB primitives in FP64 (kaiser_bessel.rs:55-164) ------------------------
double kb_i0(double z) {
  double scaled = 0.25 * z * z, sum = 1.0, term = 1.0;
  for (int k = 1; k <= 256; ++k) {
    double kf = (double)k;
    term *= scaled / (kf * kf);
    sum += term;
    if (term <= 2.220446049250313e-16 * sum) break;
  }
  return sum;
}

double kb_i1(double z) {
  // I1(z) = sum_k (z/2)^(2k+1) / (k! (k+1)!)
  double half = 0.5 * z, q = half * half, term = half, sum = half;
  for (int k = 1; k <= 256; ++k) {
    term *= q / ((double)k * (double)(k + 1));
    sum += term;
    if (std::fabs(term) <= 2.220446049250313e-16 * std::fabs(sum)) break;
  }
  return sum;
}

double kb_phi(double x, double w, double beta, double i0b) {
  double u2 = (x / w) * (x / w);
  if (u2 >= 1.0) return 0.0;
  return kb_i0(beta * std::sqrt(1.0 - u2)) / i0b;
}

double kb_dphi(double x, double w, double beta, double i0b) {
  double u2 = (x / w) * (x / w);
  if (u2 >= 1.0) return 0.0;
  double s = std::sqrt(1.0 - u2);
  if (s < 1e-300) return 0.0;
  return kb_i1(beta * s) * beta * (-x / (w * w * s)) / i0b;
}

double kb_phi_hat(double xi, int w, double beta, double i0b) {
  double a = 2.0 * M_PI * (double)w * xi;
  double z2 = beta * beta - a * a;
  double prefix = 2.0 * (double)w / i0b;
  if (std::fabs(z2) < 1e-30) return prefix;
  if (z2 > 0.0) {
    double s = std::sqrt(z2);
    return prefix * std::sinh(s) / s;
  }
  double s = std::sqrt(-z2);
  return prefix * std::sin(s) / s;
}

int64_t signed_index(int64_t k, int64_t n) { return k <= n / 2 ? k : k - n; }

// Chebyshev interpolation of f on [-1,1] at kPolyCoefs nodes, returned as
// monomial coefficients in z (FP64).
template <typename F>
void cheb_fit(F f, double *mono) {
  const int N = kPolyCoefs;
  double fz[N], a[N];
  for (int k = 0; k < N; ++k) fz[k] = f(std::cos(M_PI * (k + 0.5) / N));
  for (int j = 0; j < N; ++j) {
    double s = 0.0;
    for (int k = 0; k < N; ++k) s += fz[k] * std::cos(M_PI * j * (k + 0.5) / N);
    a[j] = s * 2.0 / N;
  }
  a[0] *= 0.5;
  // T_j in the monomial basis by T_{j+1} = 2 z T_j - T_{j-1}
  double Tm1[N] = {0}, T0[N] = {0}, T1[N] = {0};
  T0[0] = 1.0;
  for (int d = 0; d < N; ++d) mono[d] = 0.0;
  for (int j = 0; j < N; ++j) {
    for (int d = 0; d < N; ++d) mono[d] += a[j] * T0[d];
    // next: T_1 = z, then T_{j+1} = 2 z T_j - T_{j-1}
    const double mult = j == 0 ? 1.0 : 2.0;
    for (int d = 0; d < N; ++d) T1[d] = -Tm1[d] + (d > 0 ? mult * T0[d - 1] : 0.0);
    for (int d = 0; d < N; ++d) {
      Tm1[d] = T0[d];
      T0[d] = T1[d];
    }
  }
}

double horner_f32(const float c[kPolyCoefs][kMaxTaps], int tap, double z) {
  double v = c[kPolyCoefs - 1][tap];
  for (int d = kPolyCoefs - 2; d >= 0; --d) v = v * z + (double)c[d][tap];
  return v;
}

// Stockham factor order: radix 8 first, then 4/2, 3, 5, other primes.
std::vector<int> factorize(int L) {
  std::vector<int> f;
  int n = L;
  while (n % 8 == 0) {
    f.push_back(8);
    n /= 8;
  }
  if (n % 4 == 0) {
    f.push_back(4);
    n /= 4;
  }
  if (n % 2 == 0) {
    f.push_back(2);
    n /= 2;
  }
  for (int p = 3; n > 1 && p * p <= n; p += 2)
    while (n % p == 0) {
      f.push_back(p);
      n /= p;
    }
  if (n > 1) f.push_back(n);
  return f;
}

size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

// Per-stage HIP-event timing (diagnostic; enabled per plan).  The mutex
// guards the event pool and records only; it is never held while a stage
// runs, so it nests inside no other plan lock.
struct StageTimer {
  std::mutex mu;
  std::atomic<bool> on{false};
  struct Rec {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<hipEvent_t> pool;
  std::vector<Rec> recs;
  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    // timing only: no system-scope fence (cache writeback) when recorded,
    // so timing a call does not slow it down
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    return e;
  }
};

struct Layout {
  size_t keys0 = 0, keys1 = 0, idx0 = 0, idx1 = 0, cnt = 0, scan = 0, bin_start = 0,
         sub_ofs = 0, subs = 0, rec = 0, xo = 0, zrec = 0, zrec2 = 0, tmax = 0, grid = 0, fftbuf = 0,
         total = 0;
  int nchunks = 0;
};

} // namespace

struct anufft_plan {
  int dim = 3;
  int64_t n[3] = {1, 1, 1};
  double d[3] = {1, 1, 1};
  double L[3] = {1, 1, 1};
  int64_t m[3] = {1, 1, 1};
  int sigma = 2, w = 6;
  double beta = 0, i0b = 1;
  anufft_opts opts{};
  int device = 0;
  int cus = 256;
  std::vector<double> deconv[3];
  double *d_deconv[3] = {nullptr, nullptr, nullptr};
  float2 *d_tw[3] = {nullptr, nullptr, nullptr};
  std::vector<int> fac[3];          // radices of the in-LDS transform length per axis
  int blue_n[3] = {0, 0, 0};        // Bluestein length per axis (0: direct mixed radix)
  float2 *d_chirp[3] = {nullptr, nullptr, nullptr};
  float2 *d_bhat[3] = {nullptr, nullptr, nullptr};
  KbPoly *d_poly = nullptr;         // device copy of poly (read by the MFMA kernels)
  KbPoly poly{};
  double poly_err = 0;
  // Binning geometries: spreading (register columns) and interpolation /
  // gradient (LDS tile).  Each execute sorts with the geometry it uses.
  struct Binning {
    Geom g{};
    int nbins = 1;
    int passes = 1;
    int max_sub = 1024;
  } bs, bi, bg;
  int lds_max = 163840; // gfx950 LDS per workgroup
  int kernel_method = kSpreadMfma;
  hipStream_t host_stream = nullptr;
  // Second stream for the Type-2 FFT passes, which run concurrently with the
  // binning of the points (fork/join through the two events; the mutex keeps
  // concurrent callers from interleaving their record/wait pairs).
  hipStream_t aux_stream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  mutable std::mutex fork_mu;
  mutable StageTimer timer;

  int64_t modes() const { return n[0] * n[1] * n[2]; }
  int64_t cells() const { return m[0] * m[1] * m[2]; }
};

namespace {

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (dev >= 0 && dev != prev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

Layout layout(const anufft_plan *p, int64_t M) {
  Layout l;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += align_up(bytes);
    return o;
  };
  {
    int64_t nch = M > 0 ? (M + kSortChunk - 1) / kSortChunk : 1;
    l.nchunks = (int)nch;
    size_t Mp = (size_t)std::max<int64_t>(M, 1);
    l.keys0 = take(4 * Mp);
    l.keys1 = take(4 * Mp);
    l.idx0 = take(4 * Mp);
    l.idx1 = take(4 * Mp);
    const int nb = std::max(p->bs.nbins, std::max(p->bi.nbins, p->bg.nbins));
    size_t cnt_n = 0;
    for (const anufft_plan::Binning *b : {&p->bs, &p->bi, &p->bg})
      cnt_n = std::max(cnt_n, sort_count_entries(M, b->nbins));
    l.cnt = take(4 * cnt_n);
    l.scan = take(4 * (size_t)std::max(kRadix, nb));  // digit (or tile) totals
    l.bin_start = take(4 * (size_t)(nb + 1));
    l.sub_ofs = take(4 * (size_t)(nb + 2));
    int64_t nsub = 1;
    for (const anufft_plan::Binning *b : {&p->bs, &p->bi, &p->bg})
      nsub = std::max<int64_t>(nsub, b->nbins + (M + b->max_sub - 1) / b->max_sub + 1);
    l.subs = take(16 * (size_t)nsub);
    if (p->dim == 1) {
      l.zrec = take(32 * Mp);   // point records of the 1D binning
      l.zrec2 = take(32 * Mp);
      l.grid = take(8 * (size_t)p->m[0]);
      l.total = off;
      return l;
    }
    if (p->kernel_method == kSpreadMfma) {
      l.zrec = take(32 * Mp);   // ZRec: 32-byte record per point
      l.zrec2 = take(32 * Mp);  // the other record buffer of the binning passes
      l.tmax = take(2 * sizeof(float) * (size_t)(p->bi.nbins + 1));  // bin and tile maxima
    }
    else
      l.rec = take(sizeof(float) * (size_t)point_rec_floats(p->w) * Mp);
    l.xo = take(Mp);
    l.grid = take(8 * (size_t)p->cells());
    l.fftbuf = take(8 * (size_t)(p->m[0] * p->m[1] * p->n[2]));
  }
  l.total = off;
  return l;
}

hipError_t upload(float2 **dst, const std::vector<float2> &v, const char *what) {
  hipError_t e = hipMalloc(dst, sizeof(float2) * v.size());
  if (e == hipSuccess) e = hipMemcpy(*dst, v.data(), sizeof(float2) * v.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) (void)hip_fail(e, what);
  return e;
}

// In-place FP64 radix-2 DFT, exp(-2 pi i nk / n), n a power of two.
void fft_pow2_f64(std::vector<std::complex<double>> &x) {
  const size_t n = x.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(x[i], x[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    for (size_t i = 0; i < n; i += len)
      for (size_t k = 0; k < len / 2; ++k) {
        const double ang = -2.0 * M_PI * (double)k / (double)len;
        const std::complex<double> w(std::cos(ang), std::sin(ang));
        const std::complex<double> u = x[i + k], v = x[i + k + len / 2] * w;
        x[i + k] = u + v;
        x[i + k + len / 2] = u - v;
      }
  }
}

// Bluestein tables of a length-L DFT done with length-n2 FFTs:
// chirp_n = exp(-i pi n^2 / L) (n^2 reduced mod 2L in integers, so the
// angle is exact before the FP64 sine/cosine), and the spectrum of the
// filter b_m = conj(chirp_|m|) (m in (-L, L), circularly), scaled by 1/n2.
void bluestein_tables(int L, int n2, std::vector<float2> &chirp, std::vector<float2> &bhat) {
  std::vector<std::complex<double>> c(L), b(n2, 0.0);
  for (int n = 0; n < L; ++n) {
    const long long r = ((long long)n * n) % (2LL * L);
    const double ang = -M_PI * (double)r / (double)L;
    c[n] = std::complex<double>(std::cos(ang), std::sin(ang));
  }
  for (int m = 0; m < L; ++m) {
    b[m] = std::conj(c[m]);
    if (m > 0) b[n2 - m] = std::conj(c[m]);
  }
  fft_pow2_f64(b);
  chirp.resize(L);
  bhat.resize(n2);
  for (int n = 0; n < L; ++n) chirp[n] = make_float2((float)c[n].real(), (float)c[n].imag());
  for (int k = 0; k < n2; ++k)
    bhat[k] = make_float2((float)(b[k].real() / n2), (float)(b[k].imag() / n2));
}

anufft_status build_tables(anufft_plan *p) {
  // deconvolution 1/phi_hat(k_signed/m) (kaiser_bessel.rs:137-148)
  for (int a = 0; a < p->dim; ++a) {
    p->deconv[a].resize(p->n[a]);
    for (int64_t k = 0; k < p->n[a]; ++k) {
      double xi = (double)signed_index(k, p->n[a]) / (double)p->m[a];
      p->deconv[a][k] = 1.0 / kb_phi_hat(xi, p->w, p->beta, p->i0b);
    }
  }
  // per-tap polynomials of phi and d phi/dx on x = -w + i + (z+1)/2.  Tap
  // 2w-1-i at -z sits at -x, and phi is even (d phi/dx odd), so the fits of
  // taps 0 .. w-1 are mirrored: P_{2w-1-i}(z) = P_i(-z), D_{2w-1-i}(z) =
  // -D_i(-z); the kernels evaluate a tap pair from the even and odd parts.
  const int T = 2 * p->w;
  const double wd = (double)p->w;
  double mono[kPolyCoefs];
  for (int i = 0; i < p->w; ++i) {
    cheb_fit([&](double z) { return kb_phi(-wd + i + 0.5 * (z + 1.0), wd, p->beta, p->i0b); },
             mono);
    for (int d = 0; d < kPolyCoefs; ++d) {
      p->poly.c[d][i] = (float)mono[d];
      p->poly.c[d][T - 1 - i] = (float)((d & 1) ? -mono[d] : mono[d]);
    }
    cheb_fit([&](double z) { return kb_dphi(-wd + i + 0.5 * (z + 1.0), wd, p->beta, p->i0b); },
             mono);
    for (int d = 0; d < kPolyCoefs; ++d) {
      p->poly.dc[d][i] = (float)mono[d];
      p->poly.dc[d][T - 1 - i] = (float)((d & 1) ? mono[d] : -mono[d]);
    }
  }
  double err = 0.0;
  for (int i = 0; i < T; ++i)
    for (int s = 1; s < 256; ++s) {
      double z = -1.0 + 2.0 * s / 256.0;
      double x = -wd + i + 0.5 * (z + 1.0);
      err = std::max(err, std::fabs(horner_f32(p->poly.c, i, z) - kb_phi(x, wd, p->beta, p->i0b)));
    }
  p->poly_err = err;
  for (int a = 0; a < p->dim; ++a) {
    const int L = (int)p->m[a];
    p->fac[a] = factorize(L);
    const int pmax = *std::max_element(p->fac[a].begin(), p->fac[a].end());
    int n2 = 1;
    while (n2 < 2 * L - 1) n2 <<= 1;
    if (L > 1 && pmax > kDirectPrimeMax && n2 <= kBluesteinMaxN) {
      p->blue_n[a] = n2;
      p->fac[a] = factorize(n2);
    }
  }
  if (p->device < 0) return ANUFFT_OK;

  DeviceGuard guard(p->device);
  {
    hipError_t e = hipMalloc(&p->d_poly, sizeof(KbPoly));
    if (e == hipSuccess) e = hipMemcpy(p->d_poly, &p->poly, sizeof(KbPoly), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "device kernel polynomials");
  }
  hipError_t es = hipStreamCreateWithFlags(&p->host_stream, hipStreamNonBlocking);
  if (es == hipSuccess) es = hipStreamCreateWithFlags(&p->aux_stream, hipStreamNonBlocking);
  if (es == hipSuccess) es = hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming);
  if (es == hipSuccess) es = hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming);
  if (es != hipSuccess) return hip_fail(es, "hipStreamCreate");
  for (int a = 0; a < p->dim; ++a) {
    hipError_t e = hipMalloc(&p->d_deconv[a], sizeof(double) * p->n[a]);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(deconv)");
    e = hipMemcpy(p->d_deconv[a], p->deconv[a].data(), sizeof(double) * p->n[a],
                  hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(deconv)");
    const int L = (int)p->m[a];
    const int Lf = p->blue_n[a] ? p->blue_n[a] : L;
    std::vector<float2> tw(Lf);
    for (int k = 0; k < Lf; ++k) {
      double ang = -2.0 * M_PI * (double)k / (double)Lf;
      tw[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    e = upload(&p->d_tw[a], tw, "twiddles");
    if (e != hipSuccess) return ANUFFT_HIP_ERROR;
    if (p->blue_n[a]) {
      std::vector<float2> ch, bh;
      bluestein_tables(L, Lf, ch, bh);
      if (upload(&p->d_chirp[a], ch, "chirp") != hipSuccess) return ANUFFT_HIP_ERROR;
      if (upload(&p->d_bhat[a], bh, "chirp filter") != hipSuccess) return ANUFFT_HIP_ERROR;
    }
  }
  return ANUFFT_OK;
}

void destroy_plan(anufft_plan *p) {
  if (!p) return;
  if (p->device >= 0) {
    DeviceGuard guard(p->device);
    for (int a = 0; a < 3; ++a) {
      if (p->d_deconv[a]) (void)hipFree(p->d_deconv[a]);
      if (p->d_tw[a]) (void)hipFree(p->d_tw[a]);
      if (p->d_chirp[a]) (void)hipFree(p->d_chirp[a]);
      if (p->d_bhat[a]) (void)hipFree(p->d_bhat[a]);
    }
    if (p->d_poly) (void)hipFree(p->d_poly);
    if (p->host_stream) (void)hipStreamDestroy(p->host_stream);
    if (p->aux_stream) (void)hipStreamDestroy(p->aux_stream);
    if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
    if (p->ev_join) (void)hipEventDestroy(p->ev_join);
    for (auto &r : p->timer.recs) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (auto e : p->timer.pool) (void)hipEventDestroy(e);
  }
  delete p;
}

// A plan records its device; without any HIP device the plan still carries
// its KB metadata (host tables) and execute calls report
// ANUFFT_DEVICE_UNAVAILABLE.
void init_device(anufft_plan *p) {
  int count = 0;
  p->device = -1;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return;
  int dev = p->opts.device_id;
  if (dev < 0 && hipGetDevice(&dev) != hipSuccess) dev = 0;
  if (dev >= count) return;
  p->device = dev;
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
    p->cus = v;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
    p->lds_max = v;
}

// Binning geometries.  Spreading: tiles of kSpreadBinX x (32-2w) x (32-2w)
// cells, so that the padded (y,z) extent matches the 32 x 32 lane columns of
// the spreading workgroup.  Interpolation: the tile shape whose padded LDS
// copy (and the gradient kernel's) fits one workgroup's LDS and maximises
// interior / padded volume.
void set_geom(anufft_plan::Binning &b, const anufft_plan *p, const int bin[3]) {
  Geom &g = b.g;
  for (int a = 0; a < 3; ++a) {
    g.m[a] = (int)p->m[a];
    g.scale[a] = (double)p->m[a] / p->L[a];
    g.inv_m[a] = 1.0 / (double)p->m[a];
    g.bin[a] = std::max(1, std::min(bin[a], g.m[a]));
    g.nbin[a] = (g.m[a] + g.bin[a] - 1) / g.bin[a];
    g.pad[a] = g.bin[a] + 2 * p->w;
    g.inv_bin[a] = 1.0f / (float)g.bin[a];
  }
  b.nbins = g.nbin[0] * g.nbin[1] * g.nbin[2];
  int bits = 1;
  while ((1ll << bits) < (long long)b.nbins) ++bits;
  b.passes = (bits + kRadixBits - 1) / kRadixBits;
}

void choose_bins(anufft_plan *p) {
  if (p->dim == 1) {
    // tiles of one wave (64 cells with their taps); subproblems of at most
    // 1024 points (one workgroup walks a subproblem 64 points at a time)
    const int b[3] = {bin1d(p->w), 1, 1};
    set_geom(p->bs, p, b);
    p->bs.max_sub = p->opts.max_subproblem > 0 ? p->opts.max_subproblem : 1024;
    p->bi = p->bs;
    p->bg = p->bs;
    return;
  }
  const bool user = p->opts.bin_size[0] > 0 && p->opts.bin_size[1] > 0 && p->opts.bin_size[2] > 0;
  const int cyz = kSpreadCols - 2 * p->w;
  // matrix-core tiles: y and z cut into as many tiles as cyz-cell bins need,
  // with the cells shared evenly (128 cells: 8 tiles of 16 rather than 7 of
  // 18 and one of 2), so that no tile pays the fixed cost of a tile for a
  // sliver of the points
  auto even = [&](int a) {
    const int m = (int)p->m[a], nt = (m + cyz - 1) / cyz;
    return p->kernel_method == kSpreadMfma ? (m + nt - 1) / nt : cyz;
  };
  const int bsp[3] = {spread_bin_x(p->w, p->kernel_method), even(1), even(2)};
  set_geom(p->bs, p, bsp);
  p->bs.max_sub = p->opts.max_subproblem > 0 ? p->opts.max_subproblem : 4096;

  // LDS-tile geometry (gradient; interpolation on the vector-ALU path)
  const int cand[] = {32, 24, 16, 12, 10, 8, 6, 4, 3, 2, 1};
  if (user) {
    set_geom(p->bg, p, p->opts.bin_size);
  } else {
    double best = -1.0;
    int bb[3] = {1, 1, 1};
    anufft_plan::Binning t;
    for (int bx : cand)
      for (int by : cand)
        for (int bz : cand) {
          const int b[3] = {bx, by, bz};
          set_geom(t, p, b);
          size_t need = std::max(interp_lds_bytes(p->w, t.g, false), interp_lds_bytes(p->w, t.g, true));
          if (need > (size_t)p->lds_max) continue;
          double in = (double)t.g.bin[0] * t.g.bin[1] * t.g.bin[2];
          double pad = (double)t.g.pad[0] * t.g.pad[1] * t.g.pad[2];
          double score = in / pad + 1e-9 * in;
          if (score > best) {
            best = score;
            bb[0] = bx;
            bb[1] = by;
            bb[2] = bz;
          }
        }
    set_geom(p->bg, p, bb);
  }
  p->bg.max_sub = p->opts.max_subproblem > 0 ? p->opts.max_subproblem : 1024;
  if (p->kernel_method == kSpreadMfma) {
    set_geom(p->bi, p, bsp);  // MFMA interpolation uses the spreading tiles
    p->bi.max_sub = p->bs.max_sub;
  } else {
    p->bi = p->bg;
  }
}

anufft_status finish_create(anufft_plan *p, anufft_plan **out) {
  init_device(p);
  if (p->opts.device_id >= 0 && p->device != p->opts.device_id) {
    int id = p->opts.device_id;
    destroy_plan(p);
    return fail(ANUFFT_DEVICE_UNAVAILABLE, "device_id %d is not available", id);
  }
  p->beta = M_PI * (1.0 - 1.0 / (2.0 * (double)p->sigma)) * (double)(2 * p->w);
  p->i0b = kb_i0(p->beta);
  choose_bins(p);
  if (p->dim == 3) {
    size_t need = std::max(spread_lds_bytes(p->w, p->kernel_method, p->bs.g),
                           std::max(interp_lds_bytes(p->w, p->bg.g, false),
                                    interp_lds_bytes(p->w, p->bg.g, true)));
    if (need > (size_t)p->lds_max) {
      destroy_plan(p);
      return fail(ANUFFT_INVALID_ARG, "bin_size needs %zu bytes of LDS, device has %d", need,
                  p->lds_max);
    }
  }
  for (int a = 0; a < p->dim; ++a) {
    // one LDS line must hold the oversampled length (two buffers, 64 KiB)
    if (p->m[a] > 4096) {
      destroy_plan(p);
      return fail(ANUFFT_UNSUPPORTED, "oversampled length %lld exceeds the in-LDS FFT limit 4096",
                  (long long)p->m[a]);
    }
  }
  anufft_status st = build_tables(p);
  if (st != ANUFFT_OK) {
    destroy_plan(p);
    return st;
  }
  *out = p;
  return ANUFFT_OK;
}

// Launch settings of anufft_opts (speed only): 0 = automatic.
anufft_status check_launch_opts(const anufft_opts *o) {
  if (!o) return ANUFFT_OK;
  if (o->sub_work < 0 || o->sub_work > 65536 || o->blocks_per_cu < 0 || o->blocks_per_cu > 16 ||
      o->fft_lines < 0 || o->fft_lines > 4096)
    return fail(ANUFFT_INVALID_ARG,
                "launch settings out of range: sub_work %d (0..65536), blocks_per_cu %d (0..16), "
                "fft_lines %d (0..4096)", o->sub_work, o->blocks_per_cu, o->fft_lines);
  return ANUFFT_OK;
}

anufft_status check_common(int32_t sigma, int32_t w) {
  if (sigma < 2) return fail(ANUFFT_INVALID_ARG, "sigma must be >= 2 (got %d)", sigma);
  if (w < 2) return fail(ANUFFT_INVALID_ARG, "kernel_width must be >= 2 (got %d)", w);
  if (w > kMaxW)
    return fail(ANUFFT_UNSUPPORTED, "kernel_width %d exceeds the device maximum %d", w, kMaxW);
  return ANUFFT_OK;
}

int64_t over_len(int64_t n, int sigma, int w, int policy, int dim) {
  int64_t m = n * sigma;
  if (dim == 1 && policy == ANUFFT_GRID_REFERENCE) return m; // dimension_1d.rs:131
  m = std::max<int64_t>(m, 2 * w + 1);
  if (policy == ANUFFT_GRID_EXACT) return m;
  int64_t q = 1;
  while (q < m) q <<= 1; // dimension_3d.rs:48-54
  return q;
}

// ---- FFT pass construction ---------------------------------------------------
FftPass make_pass(const anufft_plan *p, int axis, const float2 *in, float2 *out, int nin,
                  int nout, int64_t na, int64_t nb, int64_t in_sa, int64_t in_sb, int64_t in_si,
                  int64_t out_sa, int64_t out_sb, int64_t out_si, bool inverse, double scale) {
  FftPass f{};
  f.in = in;
  f.out = out;
  f.L = (int)p->m[axis];
  f.nin = nin;
  f.nout = nout;
  f.nlines = na * nb;
  f.nb = nb;
  f.in_sa = in_sa;
  f.in_sb = in_sb;
  f.in_si = in_si;
  f.out_sa = out_sa;
  f.out_sb = out_sb;
  f.out_si = out_si;
  f.dec_in = inverse ? p->d_deconv[axis] : nullptr;
  f.dec_out = inverse ? nullptr : p->d_deconv[axis];
  f.scale = scale;
  f.inverse = inverse ? 1 : 0;
  f.tw = p->d_tw[axis];
  f.nfac = (int)p->fac[axis].size();
  for (int i = 0; i < f.nfac && i < 16; ++i) f.fac[i] = p->fac[axis][i];
  f.blue_n = p->blue_n[axis];
  f.chirp = p->d_chirp[axis];
  f.bhat = p->d_bhat[axis];
  // lines per block: two LDS buffers of lb*Lf complex within 16 KiB (more,
  // smaller blocks keep more loads in flight: 9 % faster than 64 KiB at
  // 128^3), and enough blocks to cover the chip when there are few lines.
  const int Lf = f.blue_n ? f.blue_n : f.L;
  int lb = std::max(1, 1024 / Lf);
  while (lb > 8 && (f.nlines + lb - 1) / lb < 2 * (int64_t)p->cus) lb >>= 1;
  // plan setting: lines per workgroup, within 64 KiB of LDS buffers
  if (p->opts.fft_lines > 0) lb = std::max(1, std::min(p->opts.fft_lines, 4096 / Lf));
  f.lb = (int)std::max<int64_t>(1, std::min<int64_t>(lb, f.nlines));
  return f;
}

struct Exec {
  const anufft_plan *p;
  hipStream_t s;
  char *ws;
  Layout lay;
  // Run f, bracketed by events on the execution stream when timing is on.
  template <typename F>
  hipError_t timed(int stage, F f) const {
    StageTimer &t = p->timer;
    if (!t.on.load()) return f();
    hipEvent_t a, b;
    {
      std::lock_guard<std::mutex> lk(t.mu);
      a = t.get();
      b = t.get();
    }
    if (a) (void)hipEventRecord(a, s);
    hipError_t e = f();
    if (b) (void)hipEventRecord(b, s);
    std::lock_guard<std::mutex> lk(t.mu);
    if (a && b) t.recs.push_back({stage, a, b});
    else {
      if (a) t.pool.push_back(a);
      if (b) t.pool.push_back(b);
    }
    return e;
  }
  // Run aux (an Exec on the plan's second stream) concurrently with main
  // (on this stream); this stream then waits for aux.  The record/wait pairs
  // are enqueued under the plan's mutex.
  template <typename A, typename B>
  hipError_t fork_join(A aux, B main) const {
    std::lock_guard<std::mutex> lk(p->fork_mu);
    Exec ea = *this;
    ea.s = p->aux_stream;
    hipError_t e = hipEventRecord(p->ev_fork, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(ea.s, p->ev_fork, 0);
    if (e == hipSuccess) e = aux(ea);
    if (e == hipSuccess) e = hipEventRecord(p->ev_join, ea.s);
    if (e != hipSuccess) return e;
    hipError_t em = main();
    hipError_t ej = hipStreamWaitEvent(s, p->ev_join, 0);
    return em != hipSuccess ? em : ej;
  }
  float2 *grid() const { return (float2 *)(ws + lay.grid); }
  float *rec() const { return (float *)(ws + lay.rec); }
  uint8_t *xo() const { return (uint8_t *)(ws + lay.xo); }
  ZRec zrec(int64_t, bool second = false) const {
    ZRec z{};
    z.r = (float4 *)(ws + (second ? lay.zrec2 : lay.zrec));
    return z;
  }
  // Binning; with records (MFMA kernels) the sort passes also build the
  // compact per-point records in sorted order.
  hipError_t bin(const anufft_plan::Binning &b, const PointArgs &pa, const float2 *values,
                 bool records) const {
    return timed(ANUFFT_STAGE_BIN, [&] {
      if (!records) return sort_points(pa, b.g, p->dim, sort_buffers(b), s);
      const RecordOut ro{zrec(pa.num_points), zrec(pa.num_points, true), values};
      return sort_points(pa, b.g, p->dim, sort_buffers(b), s, &ro);
    });
  }
  // Full weight records of sorted points for the vector-ALU spreading kernel.
  hipError_t prep(const anufft_plan::Binning &b, const PointArgs &pa, const float2 *values) const {
    return timed(ANUFFT_STAGE_PREP, [&] {
      return launch_point_prep(p->w, pa, values, b.g, p->poly, rec(), xo(), s);
    });
  }
  // Type-1 spreading of binned points (records, then the tile kernel).
  hipError_t spread3(const PointArgs &pa, const float2 *values, float2 *grid) const {
    if (p->kernel_method != kSpreadMfma) {
      hipError_t e = prep(p->bs, pa, values);
      if (e != hipSuccess) return e;
    }
    return timed(ANUFFT_STAGE_SPREAD, [&] {
      return launch_spread3d(p->w, p->kernel_method, pa, rec(), xo(), zrec(pa.num_points), grid,
                             p->bs.g, p->poly, p->d_poly, blocks(p->bs, pa.num_points, 2), s);
    });
  }
  float2 *fftbuf() const { return (float2 *)(ws + lay.fftbuf); }
  SortBuffers sort_buffers(const anufft_plan::Binning &b) const {
    SortBuffers sb{};
    sb.keys[0] = (uint32_t *)(ws + lay.keys0);
    sb.keys[1] = (uint32_t *)(ws + lay.keys1);
    sb.idx[0] = (uint32_t *)(ws + lay.idx0);
    sb.idx[1] = (uint32_t *)(ws + lay.idx1);
    sb.cnt = (uint32_t *)(ws + lay.cnt);
    sb.scan_tmp = (uint32_t *)(ws + lay.scan);
    sb.bin_start = (uint32_t *)(ws + lay.bin_start);
    sb.sub_ofs = (uint32_t *)(ws + lay.sub_ofs);
    sb.subs = (int4 *)(ws + lay.subs);
    sb.nbins = b.nbins;
    sb.nchunks = lay.nchunks;
    sb.passes = b.passes;
    sb.max_sub = b.max_sub;
    sb.sub_work = p->opts.sub_work > 0 ? p->opts.sub_work : (int)kSubWork;
    return sb;
  }
  PointArgs point_args(const anufft_plan::Binning &b, int64_t M, const float *const c[3],
                       int64_t stride) const {
    PointArgs pa{};
    pa.x = c[0];
    pa.y = p->dim == 3 ? c[1] : c[0];
    pa.z = p->dim == 3 ? c[2] : c[0];
    pa.stride = stride;
    pa.num_points = M;
    SortBuffers sb = sort_buffers(b);
    pa.perm = sb.idx[b.passes & 1];
    pa.subs = sb.subs;
    pa.nsub = sb.sub_ofs + b.nbins;
    return pa;
  }
  // Workgroups of a 3D point kernel: per_cu per CU (the plan's blocks_per_cu
  // when set), at most one per subproblem.
  int blocks(const anufft_plan::Binning &b, int64_t M, int per_cu) const {
    if (p->opts.blocks_per_cu > 0) per_cu = p->opts.blocks_per_cu;
    int64_t up = (int64_t)b.nbins + (M + b.max_sub - 1) / b.max_sub;
    int64_t n = std::min<int64_t>(up, (int64_t)per_cu * p->cus);
    return (int)std::max<int64_t>(n, 1);
  }

  // T1 grid -> modes: forward passes z, y, x with crop + deconvolution.
  hipError_t grid_to_modes(float2 *out) const {
    const int64_t mx = p->m[0], my = p->m[1], mz = p->m[2];
    const int64_t nx = p->n[0], ny = p->n[1], nz = p->n[2];
    float2 *g = grid(), *a1 = fftbuf();
    hipError_t e;
    if (p->dim == 1) {
      FftPass f = make_pass(p, 0, g, out, (int)mx, (int)nx, 1, 1, 0, 0, 1, 0, 0, 1, false, 1.0);
      return launch_fft_pass(f, s);
    }
    // z: grid[mx][my][mz] -> a1[mx][my][nz]
    FftPass fz = make_pass(p, 2, g, a1, (int)mz, (int)nz, mx, my, my * mz, mz, 1, my * nz, nz, 1,
                           false, 1.0);
    if ((e = launch_fft_pass(fz, s)) != hipSuccess) return e;
    // y: a1[mx][my][nz] -> a2[mx][ny][nz] (in the grid buffer)
    FftPass fy = make_pass(p, 1, a1, g, (int)my, (int)ny, mx, nz, my * nz, 1, nz, ny * nz, 1, nz,
                           false, 1.0);
    if ((e = launch_fft_pass(fy, s)) != hipSuccess) return e;
    // x: a2[mx][ny][nz] -> out[nx][ny][nz]
    FftPass fx = make_pass(p, 0, g, out, (int)mx, (int)nx, ny, nz, nz, 1, ny * nz, nz, 1, ny * nz,
                           false, 1.0);
    return launch_fft_pass(fx, s);
  }

  // T2 modes -> grid: deconvolve, zero pad, inverse passes x, y, z.  With
  // bin_max (3D, MFMA interpolation) the passes also leave the bin maxima
  // of the grid in the first half of the tile-maximum scratch.
  hipError_t modes_to_grid(const float2 *modes, bool compat, bool bin_max = false) const {
    const int64_t mx = p->m[0], my = p->m[1], mz = p->m[2];
    const int64_t nx = p->n[0], ny = p->n[1], nz = p->n[2];
    float2 *g = grid(), *c = fftbuf();
    hipError_t e;
    if (p->dim == 1) {
      FftPass f = make_pass(p, 0, modes, g, (int)nx, (int)mx, 1, 1, 0, 0, 1, 0, 0, 1, true, 1.0);
      return launch_fft_pass(f, s);
    }
    double scale = compat ? 1.0 / ((double)mx * (double)my * (double)mz) : 1.0;
    // x: modes[nx][ny][nz] -> b[mx][ny][nz] (in the grid buffer)
    FftPass fx = make_pass(p, 0, modes, g, (int)nx, (int)mx, ny, nz, nz, 1, ny * nz, nz, 1, ny * nz,
                           true, 1.0);
    const bool bm = bin_max && p->kernel_method == kSpreadMfma;
    float *bmax = bm ? (float *)(ws + lay.tmax) : nullptr;
    if (bm) {
      fx.zero = bmax;
      fx.zero_n = p->bi.nbins;
    }
    if ((e = launch_fft_pass(fx, s)) != hipSuccess) return e;
    // y: b[mx][ny][nz] -> c[mx][my][nz]
    FftPass fy = make_pass(p, 1, g, c, (int)ny, (int)my, mx, nz, ny * nz, 1, nz, my * nz, 1, nz,
                           true, 1.0);
    if ((e = launch_fft_pass(fy, s)) != hipSuccess) return e;
    // z: c[mx][my][nz] -> grid[mx][my][mz]
    FftPass fz = make_pass(p, 2, c, g, (int)nz, (int)mz, mx, my, my * nz, nz, 1, my * mz, mz, 1,
                           true, scale);
    if (bm) {
      fz.bmax = bmax;
      for (int a = 0; a < 3; ++a) {
        fz.bm_bin[a] = p->bi.g.bin[a];
        fz.bm_nbin[a] = p->bi.g.nbin[a];
      }
    }
    return launch_fft_pass(fz, s);
  }
};

// Smallest subproblem (points) of the fused batched spreading.
constexpr int kBatchMinSub = 4096;

// Point counts: the subproblem list and the per-point records carry point
// positions as signed 32-bit integers, so at most 2^31 - 1 points per call.
constexpr int64_t kMaxPoints = 2147483647;

anufft_status check_points(int64_t M, int64_t stride) {
  if (M < 0) return fail(ANUFFT_INVALID_ARG, "num_points must be >= 0");
  if (M > kMaxPoints)
    return fail(ANUFFT_U32_OVERFLOW, "num_points %lld exceeds the limit 2^31 - 1", (long long)M);
  if (stride < 1) return fail(ANUFFT_INVALID_ARG, "coordinate stride must be >= 1 (got %lld)", (long long)stride);
  return ANUFFT_OK;
}

anufft_status check_exec(const anufft_plan *p, int64_t M, const float *const coords[3],
                         int64_t stride, const void *a, const void *b, void *ws, size_t ws_bytes,
                         Layout &lay) {
  if (!p) return fail(ANUFFT_INVALID_ARG, "plan is NULL");
  anufft_status sp = check_points(M, stride);
  if (sp != ANUFFT_OK) return sp;
  if (M > 0 && (!coords || !coords[0] || (p->dim == 3 && (!coords[1] || !coords[2]))))
    return fail(ANUFFT_INVALID_ARG, "coordinate pointer is NULL");
  if ((M > 0 && !a) || !b) return fail(ANUFFT_INVALID_ARG, "data pointer is NULL");
  if (p->device < 0) return fail(ANUFFT_DEVICE_UNAVAILABLE, "no HIP device available for this plan");
  lay = layout(p, M);
  if (!ws || ws_bytes < lay.total)
    return fail(ANUFFT_CAPACITY_EXCEEDED, "workspace of %zu bytes is smaller than the %zu required",
                ws_bytes, lay.total);
  return ANUFFT_OK;
}

// Exact direct sums (anufft_direct_type1 / _type2).
anufft_status direct(int type, const anufft_plan *p, int64_t M, const float *const coords[3],
                     int64_t stride, const float *in, double *out, void *stream) {
  if (!p) return fail(ANUFFT_INVALID_ARG, "plan is NULL");
  anufft_status sp = check_points(M, stride);
  if (sp != ANUFFT_OK) return sp;
  if (M > 0 && (!coords || !coords[0] || (p->dim == 3 && (!coords[1] || !coords[2]))))
    return fail(ANUFFT_INVALID_ARG, "coordinate pointer is NULL");
  if (M > 0 && !in) return fail(ANUFFT_INVALID_ARG, "data pointer is NULL");
  if (!out && (type == 1 || M > 0)) return fail(ANUFFT_INVALID_ARG, "output pointer is NULL");
  if (p->device < 0) return fail(ANUFFT_DEVICE_UNAVAILABLE, "no HIP device available for this plan");
  if (type == 2 && M == 0) return ANUFFT_OK;
  DeviceGuard guard(p->device);
  PointArgs pa{};
  pa.x = M > 0 ? coords[0] : nullptr;
  pa.y = M > 0 && p->dim == 3 ? coords[1] : pa.x;
  pa.z = M > 0 && p->dim == 3 ? coords[2] : pa.x;
  pa.stride = stride;
  pa.num_points = M;
  hipError_t e = launch_direct(type, pa, p->dim, p->n, p->L, (const float2 *)in, (double2 *)out,
                               (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "direct sum");
  return ANUFFT_OK;
}

// Workspace of anufft_execute_type1_noalloc: one fixed array per GPU, held
// by the loaded code object until the process ends.  Host code never passes
// the symbol itself to a kernel (on the host it names a stand-in, not device
// memory); it uses the per-device address from hipGetSymbolAddress, looked
// up once per device with that device current.
constexpr size_t kScratchBytes = (size_t)64 << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
std::vector<void *> g_scratch_ptr;  // indexed by device id; nullptr until looked up

// Device address of `scratch` on the current device `dev`.
anufft_status scratch_address(int dev, void **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (dev < 0) return fail(ANUFFT_DEVICE_UNAVAILABLE, "no HIP device available for this plan");
  if ((size_t)dev >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!ptr) return fail(ANUFFT_HIP_ERROR, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return ANUFFT_OK;
}

} // namespace

namespace anufft {
// Error report for the other translation units of the library (autotune.cpp).
anufft_status report_error(anufft_status st, const char *msg) { return fail(st, "%s", msg); }
} // namespace anufft

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

void anufft_default_opts(anufft_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->grid_policy = ANUFFT_GRID_REFERENCE;
  o->compat_reference_type2_3d_scaling = 0;
  o->device_id = -1;
}

const char *anufft_last_error(void) { return g_last_error.c_str(); }

const char *anufft_version(void) { return "apollo-nufft-hip 0.1.0 (gfx950)"; }

anufft_status anufft_plan1d_create(int64_t n, double dx, int32_t sigma, int32_t kernel_width,
                                   const anufft_opts *opts, anufft_plan **plan) {
  if (!plan) return fail(ANUFFT_INVALID_ARG, "plan out-pointer is NULL");
  *plan = nullptr;
  // UniformDomain1D::new (grid.rs:36-48)
  if (n <= 0) return fail(ANUFFT_INVALID_ARG, "invalid NUFFT metadata field n: %lld (must be > 0)", (long long)n);
  if (!std::isfinite(dx) || dx <= 0.0)
    return fail(ANUFFT_INVALID_ARG, "invalid NUFFT metadata field dx: %g (must be finite and > 0)", dx);
  anufft_status st = check_common(sigma, kernel_width);
  if (st == ANUFFT_OK) st = check_launch_opts(opts);
  if (st != ANUFFT_OK) return st;
  if (n > (int64_t)UINT32_MAX / sigma) return fail(ANUFFT_U32_OVERFLOW, "oversampled length overflows u32");
  anufft_plan *p = new (std::nothrow) anufft_plan();
  if (!p) return fail(ANUFFT_INVALID_ARG, "out of host memory");
  if (opts) p->opts = *opts;
  else anufft_default_opts(&p->opts);
  p->dim = 1;
  p->n[0] = n;
  p->d[0] = dx;
  p->L[0] = (double)n * dx;
  p->sigma = sigma;
  p->w = kernel_width;
  p->m[0] = over_len(n, sigma, kernel_width, p->opts.grid_policy, 1);
  return finish_create(p, plan);
}

anufft_status anufft_plan3d_create(int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                                   double dz, int32_t sigma, int32_t kernel_width,
                                   const anufft_opts *opts, anufft_plan **plan) {
  if (!plan) return fail(ANUFFT_INVALID_ARG, "plan out-pointer is NULL");
  *plan = nullptr;
  // UniformGrid3D::new (grid.rs:76-110)
  const int64_t nn[3] = {nx, ny, nz};
  const double dd[3] = {dx, dy, dz};
  const char *nf[3] = {"nx", "ny", "nz"}, *df[3] = {"dx", "dy", "dz"};
  for (int a = 0; a < 3; ++a)
    if (nn[a] <= 0)
      return fail(ANUFFT_INVALID_ARG, "invalid NUFFT metadata field %s: %lld (must be > 0)", nf[a], (long long)nn[a]);
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(dd[a]) || dd[a] <= 0.0)
      return fail(ANUFFT_INVALID_ARG, "invalid NUFFT metadata field %s: %g (must be finite and > 0)", df[a], dd[a]);
  anufft_status st = check_common(sigma, kernel_width);
  if (st == ANUFFT_OK) st = check_launch_opts(opts);
  if (st != ANUFFT_OK) return st;
  anufft_plan *p = new (std::nothrow) anufft_plan();
  if (!p) return fail(ANUFFT_INVALID_ARG, "out of host memory");
  if (opts) p->opts = *opts;
  else anufft_default_opts(&p->opts);
  p->dim = 3;
  p->sigma = sigma;
  p->w = kernel_width;
  for (int a = 0; a < 3; ++a) {
    p->n[a] = nn[a];
    p->d[a] = dd[a];
    p->L[a] = (double)nn[a] * dd[a];
    p->m[a] = over_len(nn[a], sigma, kernel_width, p->opts.grid_policy, 3);
  }
  // device.rs:893-898: dimensions and the grid size must fit in u32; the
  // kernels index the grid with 32-bit cells per axis and 64-bit offsets.
  if ((double)p->m[0] * (double)p->m[1] * (double)p->m[2] > 4294967295.0) {
    delete p;
    return fail(ANUFFT_U32_OVERFLOW, "oversampled grid size exceeds u32");
  }
  const int sm = p->opts.kernel_method;
  if (sm < kSpreadAuto || sm > kSpreadMfma) {
    delete p;
    return fail(ANUFFT_INVALID_ARG, "kernel_method %d is not 0 (auto), 1 (valu) or 2 (mfma)", sm);
  }
  p->kernel_method = sm == kSpreadAuto ? kSpreadMfma : sm;
  return finish_create(p, plan);
}

anufft_status anufft_plan_destroy(anufft_plan *plan) {
  destroy_plan(plan);
  return ANUFFT_OK;
}

anufft_status anufft_plan_get_meta(const anufft_plan *p, anufft_meta *meta) {
  if (!p || !meta) return fail(ANUFFT_INVALID_ARG, "NULL argument");
  std::memset(meta, 0, sizeof(*meta));
  meta->dim = p->dim;
  for (int a = 0; a < 3; ++a) {
    meta->n[a] = p->n[a];
    meta->m[a] = p->m[a];
    meta->length[a] = a < p->dim ? p->L[a] : 0.0;
    meta->bin_size[a] = p->bs.g.bin[a];
    meta->interp_bin_size[a] = p->bi.g.bin[a];
  }
  meta->kernel_width = p->w;
  meta->sigma = p->sigma;
  meta->beta = p->beta;
  meta->i0_beta = p->i0b;
  meta->grid_policy = p->opts.grid_policy;
  meta->compat_reference_type2_3d_scaling = p->opts.compat_reference_type2_3d_scaling;
  meta->max_subproblem = p->bs.max_sub;
  meta->interp_max_subproblem = p->bi.max_sub;
  meta->poly_degree = kPolyCoefs - 1;
  meta->poly_max_err = p->poly_err;
  meta->kernel_method = p->kernel_method;
  meta->sub_work = p->opts.sub_work > 0 ? p->opts.sub_work : (int)kSubWork;
  meta->blocks_per_cu = p->opts.blocks_per_cu;
  meta->fft_lines = p->opts.fft_lines;
  return ANUFFT_OK;
}

anufft_status anufft_plan_get_deconv(const anufft_plan *p, int32_t axis, double *out_n) {
  if (!p || !out_n) return fail(ANUFFT_INVALID_ARG, "NULL argument");
  if (axis < 0 || axis >= p->dim) return fail(ANUFFT_INVALID_ARG, "axis %d out of range", axis);
  std::memcpy(out_n, p->deconv[axis].data(), sizeof(double) * p->deconv[axis].size());
  return ANUFFT_OK;
}

anufft_status anufft_workspace_size(const anufft_plan *p, int64_t max_points, size_t *bytes) {
  if (!p || !bytes) return fail(ANUFFT_INVALID_ARG, "NULL argument");
  if (max_points < 0) return fail(ANUFFT_INVALID_ARG, "max_points must be >= 0");
  *bytes = layout(p, max_points).total;
  return ANUFFT_OK;
}

anufft_status anufft_execute_type1(const anufft_plan *p, int64_t M, const float *const coords[3],
                                   int64_t stride, const float *values, float *out, void *ws,
                                   size_t ws_bytes, void *stream) {
  Layout lay;
  anufft_status st = check_exec(p, M, coords, stride, values, out, ws, ws_bytes, lay);
  if (st != ANUFFT_OK) return st;
  DeviceGuard guard(p->device);
  Exec ex{p, (hipStream_t)stream, (char *)ws, lay};
  hipError_t e = hipSuccess;
  if (M > 0) {
    PointArgs pa = ex.point_args(p->bs, M, coords, stride);
    const bool records = p->dim == 1 || p->kernel_method == kSpreadMfma;
    // zeroing the grid (second stream) is independent of the binning
    e = ex.fork_join(
        [&](const Exec &ea) {
          return ea.timed(ANUFFT_STAGE_FILL, [&] { return launch_fill(ea.grid(), p->cells(), ea.s); });
        },
        [&] { return ex.bin(p->bs, pa, (const float2 *)values, records); });
    if (e != hipSuccess) return hip_fail(e, "fill / binning");
    if (p->dim == 3) {
      e = ex.spread3(pa, (const float2 *)values, ex.grid());
    } else {
      e = ex.timed(ANUFFT_STAGE_SPREAD, [&] {
        return launch_spread1d(p->w, pa, ex.zrec(M), ex.grid(), p->bs.g, p->poly,
                               ex.blocks(p->bs, M, 8), ex.s);
      });
    }
    if (e != hipSuccess) return hip_fail(e, "spread");
  } else {
    e = ex.timed(ANUFFT_STAGE_FILL, [&] { return launch_fill(ex.grid(), p->cells(), ex.s); });
    if (e != hipSuccess) return hip_fail(e, "fill");
  }
  e = ex.timed(ANUFFT_STAGE_FFT, [&] { return ex.grid_to_modes((float2 *)out); });
  if (e != hipSuccess) return hip_fail(e, "fft");
  return ANUFFT_OK;
}

anufft_status anufft_scratch_capacity(size_t *bytes) {
  if (!bytes) return fail(ANUFFT_INVALID_ARG, "NULL argument");
  *bytes = kScratchBytes;
  return ANUFFT_OK;
}

anufft_status anufft_execute_type1_noalloc(const anufft_plan *p, int64_t M,
                                           const float *const coords[3], int64_t stride,
                                           const float *values, float *out, void *stream) {
  if (!p) return fail(ANUFFT_INVALID_ARG, "plan is NULL");
  anufft_status st = check_points(M, stride);
  if (st != ANUFFT_OK) return st;
  const size_t need = layout(p, M).total;
  if (need > kScratchBytes)
    return fail(ANUFFT_CAPACITY_EXCEEDED,
                "anufft_execute_type1_noalloc: %lld points on this plan need %zu workspace bytes, "
                "more than its fixed %zu-byte scratch; use anufft_execute_type1 with a workspace",
                (long long)M, need, kScratchBytes);
  if (p->device < 0) return fail(ANUFFT_DEVICE_UNAVAILABLE, "no HIP device available for this plan");
  void *ws = nullptr;
  {
    DeviceGuard guard(p->device);
    st = scratch_address(p->device, &ws);
  }
  if (st != ANUFFT_OK) return st;
  return anufft_execute_type1(p, M, coords, stride, values, out, ws, kScratchBytes, stream);
}

anufft_status anufft_execute_type2(const anufft_plan *p, int64_t M, const float *const coords[3],
                                   int64_t stride, const float *modes, float *out, void *ws,
                                   size_t ws_bytes, void *stream) {
  Layout lay;
  anufft_status st = check_exec(p, M, coords, stride, out, modes, ws, ws_bytes, lay);
  if (st != ANUFFT_OK) return st;
  if (M == 0) return ANUFFT_OK;
  DeviceGuard guard(p->device);
  Exec ex{p, (hipStream_t)stream, (char *)ws, lay};
  hipError_t e;
  if (p->dim == 3) {
    PointArgs pa = ex.point_args(p->bi, M, coords, stride);
    // the FFT passes (modes -> grid) and the binning (points) are independent:
    // the FFT goes to the plan's second stream, joined before interpolation
    e = ex.fork_join(
        [&](const Exec &ea) {
          hipError_t r = ea.timed(ANUFFT_STAGE_FFT, [&] {
            return ea.modes_to_grid((const float2 *)modes,
                                    p->opts.compat_reference_type2_3d_scaling != 0, true);
          });
          if (r != hipSuccess) return r;
          return ea.timed(ANUFFT_STAGE_TILE_MAX, [&] {
            return launch_tile_max3d(p->w, p->kernel_method, ea.grid(),
                                     (float *)(ea.ws + lay.tmax), p->bi.g, ea.s, true);
          });
        },
        [&] { return ex.bin(p->bi, pa, nullptr, p->kernel_method == kSpreadMfma); });
    if (e != hipSuccess) return hip_fail(e, "fft / binning");
    e = ex.timed(ANUFFT_STAGE_INTERP, [&] {
      return launch_interp3d(p->w, p->kernel_method, pa, ex.zrec(M), ex.grid(),
                             (float *)(ex.ws + lay.tmax), (float2 *)out, p->bi.g, p->poly, p->d_poly,
                             ex.blocks(p->bi, M, interp_blocks_per_cu(p->w, p->kernel_method)),
                             p->cus, ex.s);
    });
  } else {
    PointArgs pa = ex.point_args(p->bi, M, coords, stride);
    e = ex.fork_join(
        [&](const Exec &ea) {
          return ea.timed(ANUFFT_STAGE_FFT, [&] { return ea.modes_to_grid((const float2 *)modes, false); });
        },
        [&] { return ex.bin(p->bi, pa, nullptr, true); });
    if (e != hipSuccess) return hip_fail(e, "fft / binning");
    e = ex.timed(ANUFFT_STAGE_INTERP, [&] {
      return launch_interp1d(p->w, pa, ex.zrec(M), ex.grid(), (float2 *)out, nullptr, p->bi.g,
                             p->poly, ex.blocks(p->bi, M, 8), ex.s);
    });
  }
  if (e != hipSuccess) return hip_fail(e, "interp");
  return ANUFFT_OK;
}

anufft_status anufft_execute_grad_points(const anufft_plan *p, int64_t M,
                                         const float *const coords[3], int64_t stride,
                                         const float *u, const float *modes_h,
                                         float *const grad[3], void *ws, size_t ws_bytes,
                                         void *stream) {
  Layout lay;
  anufft_status st = check_exec(p, M, coords, stride, u, modes_h, ws, ws_bytes, lay);
  if (st != ANUFFT_OK) return st;
  if (!grad || !grad[0]) return fail(ANUFFT_INVALID_ARG, "grad[0] is NULL");
  if (M == 0) return ANUFFT_OK;
  DeviceGuard guard(p->device);
  Exec ex{p, (hipStream_t)stream, (char *)ws, lay};
  hipError_t e;
  // the derivative always uses the unnormalized (direct-sum) scaling
  if (p->dim == 3) {
    const bool gm = grad_uses_mfma(p->w, p->kernel_method);
    const anufft_plan::Binning &b = gm ? p->bi : p->bg;
    PointArgs pa = ex.point_args(b, M, coords, stride);
    // FFT passes (second stream) concurrent with the binning
    e = ex.fork_join(
        [&](const Exec &ea) {
          return ea.timed(ANUFFT_STAGE_FFT, [&] { return ea.modes_to_grid((const float2 *)modes_h, false); });
        },
        [&] { return ex.bin(b, pa, (const float2 *)u, gm); });
    if (e != hipSuccess) return hip_fail(e, "fft / binning");
    e = ex.timed(ANUFFT_STAGE_GRAD, [&] {
      return launch_grad3d(p->w, p->kernel_method, pa, ex.zrec(M), ex.grid(), (const float2 *)u,
                           grad[0], grad[1], grad[2], b.g, p->poly, p->d_poly,
                           ex.blocks(b, M, gm ? 1 : 4), ex.s);
    });
  } else {
    // records carry u (value fields): the gradient reads it in sorted order
    PointArgs pa = ex.point_args(p->bg, M, coords, stride);
    e = ex.fork_join(
        [&](const Exec &ea) {
          return ea.timed(ANUFFT_STAGE_FFT, [&] { return ea.modes_to_grid((const float2 *)modes_h, false); });
        },
        [&] { return ex.bin(p->bg, pa, (const float2 *)u, true); });
    if (e != hipSuccess) return hip_fail(e, "fft / binning");
    e = ex.timed(ANUFFT_STAGE_GRAD, [&] {
      return launch_interp1d(p->w, pa, ex.zrec(M), ex.grid(), nullptr, grad[0], p->bg.g, p->poly,
                             ex.blocks(p->bg, M, 8), ex.s);
    });
  }
  if (e != hipSuccess) return hip_fail(e, "grad");
  return ANUFFT_OK;
}

anufft_status anufft_execute_type2_backward(const anufft_plan *p, int64_t M,
                                            const float *const coords[3], int64_t stride,
                                            const float *g, const float *modes,
                                            float *grad_modes, float *const grad_x[3], void *ws,
                                            size_t ws_bytes, void *stream) {
  Layout lay;
  anufft_status st = check_exec(p, M, coords, stride, g, modes, ws, ws_bytes, lay);
  if (st != ANUFFT_OK) return st;
  if (!grad_modes) return fail(ANUFFT_INVALID_ARG, "grad_modes is NULL");
  if (M > 0 && (!grad_x || !grad_x[0])) return fail(ANUFFT_INVALID_ARG, "grad_x[0] is NULL");
  if (p->dim != 3 || !grad_uses_mfma(p->w, p->kernel_method))
    return fail(ANUFFT_UNSUPPORTED, "the fused Type-2 backward needs a 3D plan on the MFMA kernels (w <= 7)");
  if (ws_bytes < 2 * lay.total)
    return fail(ANUFFT_CAPACITY_EXCEEDED, "workspace of %zu bytes is smaller than the %zu required",
                ws_bytes, 2 * lay.total);
  if (M == 0)  // T1 of no points is zero; no coordinates to differentiate
    return anufft_execute_type1(p, 0, coords, stride, g, grad_modes, ws, ws_bytes, stream);
  DeviceGuard guard(p->device);
  Exec ex{p, (hipStream_t)stream, (char *)ws, lay};
  // the second half of the workspace holds the fine grid of the modes
  const Exec eh{p, (hipStream_t)stream, (char *)ws + lay.total, lay};
  // the MFMA gradient bins with the spreading tiles, so one binning (with g
  // in the records) serves the spreading and the gradient kernel
  PointArgs pa = ex.point_args(p->bs, M, coords, stride);
  hipError_t e = ex.fork_join(
      [&](const Exec &ea) {
        Exec eb = eh;
        eb.s = ea.s;
        hipError_t r = eb.timed(ANUFFT_STAGE_FFT, [&] { return eb.modes_to_grid((const float2 *)modes, false); });
        if (r == hipSuccess)
          r = ea.timed(ANUFFT_STAGE_FILL, [&] { return launch_fill(ea.grid(), p->cells(), ea.s); });
        return r;
      },
      [&] { return ex.bin(p->bs, pa, (const float2 *)g, true); });
  if (e != hipSuccess) return hip_fail(e, "fft / fill / binning");
  e = ex.spread3(pa, (const float2 *)g, ex.grid());
  if (e != hipSuccess) return hip_fail(e, "spread");
  e = ex.timed(ANUFFT_STAGE_FFT, [&] { return ex.grid_to_modes((float2 *)grad_modes); });
  if (e != hipSuccess) return hip_fail(e, "fft");
  e = ex.timed(ANUFFT_STAGE_GRAD, [&] {
    return launch_grad3d(p->w, p->kernel_method, pa, ex.zrec(M), eh.grid(), (const float2 *)g,
                         grad_x[0], grad_x[1], grad_x[2], p->bi.g, p->poly, p->d_poly,
                         ex.blocks(p->bi, M, 1), ex.s);
  });
  if (e != hipSuccess) return hip_fail(e, "grad");
  return ANUFFT_OK;
}

anufft_status anufft_execute_type1_backward(const anufft_plan *p, int64_t M,
                                            const float *const coords[3], int64_t stride,
                                            const float *values, const float *G,
                                            float *grad_values, float *const grad_x[3], void *ws,
                                            size_t ws_bytes, void *stream) {
  Layout lay;
  anufft_status st = check_exec(p, M, coords, stride, values, G, ws, ws_bytes, lay);
  if (st != ANUFFT_OK) return st;
  if (M > 0 && !grad_values) return fail(ANUFFT_INVALID_ARG, "grad_values is NULL");
  if (M > 0 && (!grad_x || !grad_x[0])) return fail(ANUFFT_INVALID_ARG, "grad_x[0] is NULL");
  if (p->dim != 3 || !grad_uses_mfma(p->w, p->kernel_method))
    return fail(ANUFFT_UNSUPPORTED, "the fused Type-1 backward needs a 3D plan on the MFMA kernels (w <= 7)");
  if (M == 0) return ANUFFT_OK;
  DeviceGuard guard(p->device);
  Exec ex{p, (hipStream_t)stream, (char *)ws, lay};
  // T2(G) and the gradient both read the fine grid of G and bin the points
  // with the same tiles: one FFT (second stream) and one binning carrying
  // the values c (ignored by the interpolation, u of the gradient)
  PointArgs pa = ex.point_args(p->bi, M, coords, stride);
  hipError_t e = ex.fork_join(
      [&](const Exec &ea) {
        hipError_t r = ea.timed(ANUFFT_STAGE_FFT, [&] { return ea.modes_to_grid((const float2 *)G, false, true); });
        if (r != hipSuccess) return r;
        return ea.timed(ANUFFT_STAGE_TILE_MAX, [&] {
          return launch_tile_max3d(p->w, p->kernel_method, ea.grid(), (float *)(ea.ws + lay.tmax),
                                   p->bi.g, ea.s, true);
        });
      },
      [&] { return ex.bin(p->bi, pa, (const float2 *)values, true); });
  if (e != hipSuccess) return hip_fail(e, "fft / binning");
  e = ex.timed(ANUFFT_STAGE_INTERP, [&] {
    return launch_interp3d(p->w, p->kernel_method, pa, ex.zrec(M), ex.grid(),
                           (float *)(ex.ws + lay.tmax), (float2 *)grad_values, p->bi.g, p->poly,
                           p->d_poly, ex.blocks(p->bi, M, interp_blocks_per_cu(p->w, p->kernel_method)),
                           p->cus, ex.s);
  });
  if (e != hipSuccess) return hip_fail(e, "interp");
  e = ex.timed(ANUFFT_STAGE_GRAD, [&] {
    return launch_grad3d(p->w, p->kernel_method, pa, ex.zrec(M), ex.grid(), (const float2 *)values,
                         grad_x[0], grad_x[1], grad_x[2], p->bi.g, p->poly, p->d_poly,
                         ex.blocks(p->bi, M, 1), ex.s);
  });
  if (e != hipSuccess) return hip_fail(e, "grad");
  return ANUFFT_OK;
}

anufft_status anufft_timing_enable(const anufft_plan *p, int32_t on) {
  if (!p) return fail(ANUFFT_INVALID_ARG, "plan is NULL");
  std::lock_guard<std::mutex> lk(p->timer.mu);
  for (auto &r : p->timer.recs) {
    p->timer.pool.push_back(r.a);
    p->timer.pool.push_back(r.b);
  }
  p->timer.recs.clear();
  p->timer.on.store(on != 0);
  return ANUFFT_OK;
}

anufft_status anufft_timing_collect(const anufft_plan *p, double *ms, int64_t *calls) {
  if (!p || !ms || !calls) return fail(ANUFFT_INVALID_ARG, "NULL argument");
  std::lock_guard<std::mutex> lk(p->timer.mu);
  for (int i = 0; i < ANUFFT_STAGE_COUNT; ++i) {
    ms[i] = 0.0;
    calls[i] = 0;
  }
  hipError_t err = hipSuccess;
  for (auto &r : p->timer.recs) {
    float t = 0.f;
    hipError_t e = hipEventSynchronize(r.b);
    if (e == hipSuccess) e = hipEventElapsedTime(&t, r.a, r.b);
    if (e != hipSuccess) err = e;
    else if (r.stage >= 0 && r.stage < ANUFFT_STAGE_COUNT) {
      ms[r.stage] += (double)t;
      calls[r.stage] += 1;
    }
    p->timer.pool.push_back(r.a);
    p->timer.pool.push_back(r.b);
  }
  p->timer.recs.clear();
  if (err != hipSuccess) return hip_fail(err, "timing_collect");
  return ANUFFT_OK;
}

anufft_status anufft_direct_type1(const anufft_plan *p, int64_t M, const float *const coords[3],
                                  int64_t stride, const float *values, double *out, void *stream) {
  return direct(1, p, M, coords, stride, values, out, stream);
}

anufft_status anufft_direct_type2(const anufft_plan *p, int64_t M, const float *const coords[3],
                                  int64_t stride, const float *modes, double *out, void *stream) {
  return direct(2, p, M, coords, stride, modes, out, stream);
}

anufft_status anufft_direct_type1_select(const anufft_plan *p, int64_t M,
                                         const float *const coords[3], int64_t stride,
                                         const double *values, int64_t num_modes,
                                         const int64_t *mode_index, double *out, void *stream) {
  if (!p) return fail(ANUFFT_INVALID_ARG, "plan is NULL");
  anufft_status sp = check_points(M, stride);
  if (sp != ANUFFT_OK) return sp;
  if (M > 0 && (!coords || !coords[0] || (p->dim == 3 && (!coords[1] || !coords[2]))))
    return fail(ANUFFT_INVALID_ARG, "coordinate pointer is NULL");
  if (num_modes < 0) return fail(ANUFFT_INVALID_ARG, "num_modes must be >= 0");
  if ((M > 0 && !values) || (num_modes > 0 && (!mode_index || !out)))
    return fail(ANUFFT_INVALID_ARG, "data pointer is NULL");
  if (p->device < 0) return fail(ANUFFT_DEVICE_UNAVAILABLE, "no HIP device available for this plan");
  DeviceGuard guard(p->device);
  PointArgs pa{};
  pa.x = M > 0 ? coords[0] : nullptr;
  pa.y = M > 0 && p->dim == 3 ? coords[1] : pa.x;
  pa.z = M > 0 && p->dim == 3 ? coords[2] : pa.x;
  pa.stride = stride;
  pa.num_points = M;
  hipError_t e = launch_direct_select(pa, p->dim, p->n, p->L, (const double2 *)values, mode_index,
                                      num_modes, (double2 *)out, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "direct sum (selected modes)");
  return ANUFFT_OK;
}

anufft_status anufft_debug_num_bins(const anufft_plan *p, int64_t *num_bins) {
  if (!p || !num_bins) return fail(ANUFFT_INVALID_ARG, "NULL argument");
  *num_bins = p->bs.nbins;
  return ANUFFT_OK;
}

anufft_status anufft_debug_sort(const anufft_plan *p, int64_t M, const float *const coords[3],
                                int64_t stride, uint32_t *perm, uint32_t *bin_start, void *ws,
                                size_t ws_bytes, void *stream) {
  Layout lay;
  anufft_status st = check_exec(p, M, coords, stride, perm, bin_start, ws, ws_bytes, lay);
  if (st != ANUFFT_OK) return st;
  DeviceGuard guard(p->device);
  Exec ex{p, (hipStream_t)stream, (char *)ws, lay};
  PointArgs pa = ex.point_args(p->bs, M, coords, stride);
  SortBuffers sb = ex.sort_buffers(p->bs);
  hipError_t e = sort_points(pa, p->bs.g, p->dim, sb, ex.s);
  if (e != hipSuccess) return hip_fail(e, "binning");
  if (M > 0 && (e = hipMemcpyAsync(perm, pa.perm, 4 * (size_t)M, hipMemcpyDeviceToDevice, ex.s)) != hipSuccess)
    return hip_fail(e, "copy perm");
  if ((e = hipMemcpyAsync(bin_start, sb.bin_start, 4 * (size_t)(p->bs.nbins + 1), hipMemcpyDeviceToDevice, ex.s)) != hipSuccess)
    return hip_fail(e, "copy bin_start");
  return ANUFFT_OK;
}

anufft_status anufft_debug_spread(const anufft_plan *p, int64_t M, const float *const coords[3],
                                  int64_t stride, const float *values, float *grid, void *ws,
                                  size_t ws_bytes, void *stream) {
  Layout lay;
  anufft_status st = check_exec(p, M, coords, stride, values, grid, ws, ws_bytes, lay);
  if (st != ANUFFT_OK) return st;
  DeviceGuard guard(p->device);
  Exec ex{p, (hipStream_t)stream, (char *)ws, lay};
  hipError_t e = launch_fill((float2 *)grid, p->cells(), ex.s);
  if (e != hipSuccess) return hip_fail(e, "fill");
  if (M == 0) return ANUFFT_OK;
  PointArgs pa = ex.point_args(p->bs, M, coords, stride);
  e = ex.bin(p->bs, pa, (const float2 *)values, p->dim == 1 || p->kernel_method == kSpreadMfma);
  if (e != hipSuccess) return hip_fail(e, "binning");
  if (p->dim == 3)
    e = ex.spread3(pa, (const float2 *)values, (float2 *)grid);
  else
    e = launch_spread1d(p->w, pa, ex.zrec(M), (float2 *)grid, p->bs.g, p->poly,
                        ex.blocks(p->bs, M, 8), ex.s);
  if (e != hipSuccess) return hip_fail(e, "spread");
  return ANUFFT_OK;
}

anufft_status anufft_batch_workspace_size(const anufft_plan *const *plans, int32_t nprob,
                                          const int64_t *max_points, size_t *bytes) {
  if (!plans || !max_points || !bytes || nprob < 0) return fail(ANUFFT_INVALID_ARG, "NULL argument");
  size_t tot = 0;
  for (int i = 0; i < nprob; ++i) {
    if (!plans[i]) return fail(ANUFFT_INVALID_ARG, "plan %d is NULL", i);
    if (plans[i]->dim != 3) return fail(ANUFFT_INVALID_ARG, "batched execution takes 3D plans");
    if (max_points[i] < 0) return fail(ANUFFT_INVALID_ARG, "max_points[%d] < 0", i);
    tot += align_up(layout(plans[i], max_points[i]).total);
  }
  tot += align_up(sizeof(SpreadProblem) * (size_t)std::max(nprob, 1));
  tot += align_up(sizeof(uint32_t) * (size_t)(2 * nprob + 2));  // subproblem and work prefixes
  tot += align_up(sizeof(BinProblem) * (size_t)std::max(nprob, 1));  // fused binning table
  *bytes = tot;
  return ANUFFT_OK;
}

// Batched Type-1: the binning of all problems in shared launches (beside the
// grids' zeroing on the second stream), the spreading of every problem in
// one persistent launch, and per-problem FFT passes split between the two
// streams (problems must share kernel_width, sigma and the kernel method,
// hence the kernel polynomials; otherwise the problems run one after
// another).
anufft_status anufft_execute_type1_batched(const anufft_plan *const *plans, int32_t nprob,
                                           const int64_t *num_points, const float *const *coords,
                                           int64_t stride, const float *const *values,
                                           float *const *out, void *ws, size_t ws_bytes,
                                           void *stream) {
  size_t need = 0;
  anufft_status st = anufft_batch_workspace_size(plans, nprob, num_points, &need);
  if (st != ANUFFT_OK) return st;
  if (!ws || ws_bytes < need)
    return fail(ANUFFT_CAPACITY_EXCEEDED, "workspace of %zu bytes is smaller than the %zu required", ws_bytes, need);
  if (nprob == 0) return ANUFFT_OK;
  if (!coords || !values || !out) return fail(ANUFFT_INVALID_ARG, "NULL argument");
  bool fused = true;
  for (int i = 0; i < nprob; ++i) {
    const anufft_status sp = check_points(num_points[i], stride);
    if (sp != ANUFFT_OK) return fail(sp, "problem %d: %s", i, std::string(g_last_error).c_str());
    if (plans[i]->device < 0) return fail(ANUFFT_DEVICE_UNAVAILABLE, "no HIP device available for plan %d", i);
    if (plans[i]->device != plans[0]->device) return fail(ANUFFT_INVALID_ARG, "plans live on different devices");
    if (!out[i] || (num_points[i] > 0 && (!values[i] || !coords[3 * i] || !coords[3 * i + 1] || !coords[3 * i + 2])))
      return fail(ANUFFT_INVALID_ARG, "NULL pointer in problem %d", i);
    if (plans[i]->w != plans[0]->w || plans[i]->sigma != plans[0]->sigma ||
        plans[i]->kernel_method != plans[0]->kernel_method)
      fused = false;
  }
  const anufft_plan *p0 = plans[0];
  DeviceGuard guard(p0->device);
  hipStream_t s = (hipStream_t)stream;
  if (!fused) {
    size_t off = 0;
    for (int i = 0; i < nprob; ++i) {
      const float *c3[3] = {coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]};
      size_t sz = align_up(layout(plans[i], num_points[i]).total);
      st = anufft_execute_type1(plans[i], num_points[i], c3, stride, values[i], out[i],
                                (char *)ws + off, sz, stream);
      if (st != ANUFFT_OK) return st;
      off += sz;
    }
    return ANUFFT_OK;
  }
  std::vector<Exec> ex;
  std::vector<SpreadProblem> probs(nprob);
  ex.reserve(nprob);
  size_t off = 0;
  int64_t sub_upper = 0;
  for (int i = 0; i < nprob; ++i) {
    Layout lay = layout(plans[i], num_points[i]);
    ex.push_back(Exec{plans[i], s, (char *)ws + off, lay});
    off += align_up(lay.total);
    const float *c3[3] = {coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]};
    probs[i].pa = ex[i].point_args(plans[i]->bs, num_points[i], c3, stride);
    probs[i].rec = ex[i].rec();
    probs[i].xo = ex[i].xo();
    probs[i].zr = ex[i].zrec(num_points[i]);
    probs[i].grid = ex[i].grid();
    probs[i].g = plans[i]->bs.g;
    sub_upper += plans[i]->bs.nbins + (num_points[i] + plans[i]->bs.max_sub - 1) / plans[i]->bs.max_sub;
  }
  SpreadProblem *table = (SpreadProblem *)((char *)ws + off);
  off += align_up(sizeof(SpreadProblem) * (size_t)nprob);
  uint32_t *prefix = (uint32_t *)((char *)ws + off);
  off += align_up(sizeof(uint32_t) * (size_t)(2 * nprob + 2));
  BinProblem *bin_table = (BinProblem *)((char *)ws + off);
  hipError_t e;
  const Exec &e0 = ex[0];
  const bool mf = p0->kernel_method == kSpreadMfma;
  // the grids' zeroing and the problems' binnings are independent: the
  // zeroing and the odd problems on the first plan's second stream, even
  // problems on the caller's (small problems are launch-bound, so two
  // streams keep more of them in flight)
  // one persistent launch splits the problems' combined work prefix, so
  // every problem weighs its subproblems alike, with the automatic weight
  // (the plans' own settings are tuned for single calls)
  const int sub_work = (int)kSubWork;
  auto sort_range = [&](int parity, hipStream_t st) {
    for (int i = parity; i < nprob; i += 2) {
      const RecordOut ro{probs[i].zr, ex[i].zrec(num_points[i], true), (const float2 *)values[i]};
      SortBuffers sb = ex[i].sort_buffers(plans[i]->bs);
      sb.sub_work = sub_work;
      // the persistent launch balances work across all problems itself:
      // smaller subproblems only add fixed costs there
      sb.max_sub = std::max(sb.max_sub, kBatchMinSub);
      hipError_t r = sort_points(probs[i].pa, plans[i]->bs.g, 3, sb, st, mf ? &ro : nullptr);
      if (r != hipSuccess) return r;
    }
    return hipSuccess;
  };
  // all problems with one-pass tile counts and records: one fused binning
  // (four launches for all problems) beside the grids' zeroing
  bool one_launch = mf;
  for (int i = 0; i < nprob && one_launch; ++i) one_launch = sort_one_pass(plans[i]->bs.nbins);
  std::vector<BinProblem> bprobs;
  if (one_launch) {
    bprobs.resize(nprob);
    for (int i = 0; i < nprob; ++i) {
      const SortBuffers sb = ex[i].sort_buffers(plans[i]->bs);
      BinProblem &b = bprobs[i];
      b.pa = probs[i].pa;
      b.g = plans[i]->bs.g;
      b.cnt = sb.cnt;
      b.tot = sb.scan_tmp;
      b.bin_start = sb.bin_start;
      b.nsub = sb.sub_ofs + sb.nbins;
      b.subs = sb.subs;
      b.out = probs[i].zr;
      b.values = (const float2 *)values[i];
      b.dim = 3;
      b.nbins = sb.nbins;
      b.max_sub = std::max(sb.max_sub, kBatchMinSub);
      b.sub_work = sub_work;
    }
  }
  e = e0.timed(ANUFFT_STAGE_BIN, [&] {
    return e0.fork_join(
        [&](const Exec &ea) {
          for (int i = 0; i < nprob; ++i) {
            hipError_t r = launch_fill(ex[i].grid(), plans[i]->cells(), ea.s);
            if (r != hipSuccess) return r;
          }
          return one_launch ? hipSuccess : sort_range(1, ea.s);
        },
        [&] {
          return one_launch ? sort_points_batched(bprobs.data(), nprob, bin_table, s)
                            : sort_range(0, s);
        });
  });
  if (e != hipSuccess) return hip_fail(e, "fill / binning");
  if (!mf) {
    e = e0.timed(ANUFFT_STAGE_PREP, [&] {
      for (int i = 0; i < nprob; ++i) {
        hipError_t r = launch_point_prep(p0->w, probs[i].pa, (const float2 *)values[i],
                                         plans[i]->bs.g, plans[i]->poly, ex[i].rec(), ex[i].xo(), s);
        if (r != hipSuccess) return r;
      }
      return hipSuccess;
    });
    if (e != hipSuccess) return hip_fail(e, "spread records");
  }
  const int nblk = (int)std::max<int64_t>(1, std::min<int64_t>(sub_upper, 2 * (int64_t)p0->cus));
  e = e0.timed(ANUFFT_STAGE_SPREAD, [&] {
    return launch_spread3d_batched(p0->w, p0->kernel_method, probs.data(), nprob, table, prefix,
                                   p0->poly, p0->d_poly, nblk, s);
  });
  if (e != hipSuccess) return hip_fail(e, "spread");
  // the problems' FFT passes, odd problems on the second stream
  auto fft_range = [&](int parity, hipStream_t st) {
    for (int i = parity; i < nprob; i += 2) {
      Exec ei = ex[i];
      ei.s = st;
      hipError_t r = ei.grid_to_modes((float2 *)out[i]);
      if (r != hipSuccess) return r;
    }
    return hipSuccess;
  };
  e = e0.timed(ANUFFT_STAGE_FFT, [&] {
    return e0.fork_join([&](const Exec &ea) { return fft_range(1, ea.s); },
                        [&] { return fft_range(0, s); });
  });
  if (e != hipSuccess) return hip_fail(e, "fft");
  return ANUFFT_OK;
}

} // extern "C"

// ---- host-pointer variants ---------------------------------------------------
namespace {

uint16_t f32_to_f16(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  int32_t exp = (int32_t)((x >> 23) & 0xff) - 127 + 15;
  uint32_t mant = x & 0x7fffffu;
  if (((x >> 23) & 0xff) == 0xff) return (uint16_t)(sign | 0x7c00u | (mant ? 0x200u : 0u));
  if (exp >= 31) return (uint16_t)(sign | 0x7c00u);
  if (exp <= 0) {
    if (exp < -10) return (uint16_t)sign;
    mant |= 0x800000u;
    uint32_t shift = (uint32_t)(14 - exp);
    uint32_t half = mant >> shift;
    uint32_t rem = mant & ((1u << shift) - 1u);
    uint32_t mid = 1u << (shift - 1);
    if (rem > mid || (rem == mid && (half & 1u))) ++half;
    return (uint16_t)(sign | half);
  }
  uint32_t half = sign | ((uint32_t)exp << 10) | (mant >> 13);
  uint32_t rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (half & 1u))) ++half;
  return (uint16_t)half;
}

float f16_to_f32(uint16_t h) {
  uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  uint32_t exp = (h >> 10) & 0x1f, mant = h & 0x3ffu, x;
  if (exp == 0) {
    if (mant == 0) x = sign;
    else {
      int e = -1;
      do {
        ++e;
        mant <<= 1;
      } while (!(mant & 0x400u));
      x = sign | ((uint32_t)(127 - 15 - e) << 23) | ((mant & 0x3ffu) << 13);
    }
  } else if (exp == 31) {
    x = sign | 0x7f800000u | (mant << 13);
  } else {
    x = sign | ((exp + 127 - 15) << 23) | (mant << 13);
  }
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}

void to_c64(int storage, const void *src, int64_t n, std::vector<float2> &dst) {
  dst.resize((size_t)std::max<int64_t>(n, 1));
  for (int64_t i = 0; i < n; ++i) {
    if (storage == ANUFFT_STORAGE_C128) {
      const double *s = (const double *)src;
      dst[i] = make_float2((float)s[2 * i], (float)s[2 * i + 1]);
    } else if (storage == ANUFFT_STORAGE_C64) {
      const float *s = (const float *)src;
      dst[i] = make_float2(s[2 * i], s[2 * i + 1]);
    } else {
      const uint16_t *s = (const uint16_t *)src;
      dst[i] = make_float2(f16_to_f32(s[2 * i]), f16_to_f32(s[2 * i + 1]));
    }
  }
}

void from_c64(int storage, const std::vector<float2> &src, int64_t n, void *dst) {
  for (int64_t i = 0; i < n; ++i) {
    if (storage == ANUFFT_STORAGE_C128) {
      double *d = (double *)dst;
      d[2 * i] = src[i].x;
      d[2 * i + 1] = src[i].y;
    } else if (storage == ANUFFT_STORAGE_C64) {
      float *d = (float *)dst;
      d[2 * i] = src[i].x;
      d[2 * i + 1] = src[i].y;
    } else {
      uint16_t *d = (uint16_t *)dst;
      d[2 * i] = f32_to_f16(src[i].x);
      d[2 * i + 1] = f32_to_f16(src[i].y);
    }
  }
}

anufft_status host_exec(const anufft_plan *p, int type, int32_t storage, int32_t profile,
                        int64_t M, const double *const coords[3], int64_t stride,
                        const void *in, void *out) {
  if (!p) return fail(ANUFFT_INVALID_ARG, "plan is NULL");
  if (storage < 0 || storage > 2) return fail(ANUFFT_INVALID_ARG, "unknown storage %d", storage);
  if (profile != storage)
    return fail(ANUFFT_PROFILE_MISMATCH,
                "invalid field precision_profile: profile %d does not match storage %d", profile,
                storage);
  anufft_status sp = check_points(M, stride);
  if (sp != ANUFFT_OK) return sp;
  if (M > 0 && (!coords || !coords[0] || (p->dim == 3 && (!coords[1] || !coords[2]))))
    return fail(ANUFFT_INVALID_ARG, "coordinate pointer is NULL");
  if ((M > 0 || type == 1) && (!in || !out)) return fail(ANUFFT_INVALID_ARG, "data pointer is NULL");
  if (p->device < 0) return fail(ANUFFT_DEVICE_UNAVAILABLE, "no HIP device available for this plan");
  DeviceGuard guard(p->device);
  hipStream_t s = p->host_stream;
  const int64_t N = p->modes();
  const int64_t nin = type == 1 ? M : N, nout = type == 1 ? N : M;
  std::vector<float2> hin, hout((size_t)std::max<int64_t>(nout, 1));
  to_c64(storage, in, nin, hin);
  std::vector<float> hc((size_t)std::max<int64_t>(M, 1) * p->dim);
  for (int a = 0; a < p->dim; ++a)
    for (int64_t j = 0; j < M; ++j) hc[(size_t)a * M + j] = (float)coords[a][j * stride];
  size_t wsb = layout(p, M).total;
  char *buf = nullptr;
  size_t b_in = align_up(8 * (size_t)std::max<int64_t>(nin, 1));
  size_t b_out = align_up(8 * (size_t)std::max<int64_t>(nout, 1));
  size_t b_c = align_up(4 * hc.size());
  hipError_t e = hipMalloc(&buf, b_in + b_out + b_c + wsb);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  float2 *d_in = (float2 *)buf, *d_out = (float2 *)(buf + b_in);
  float *d_c = (float *)(buf + b_in + b_out);
  void *d_ws = buf + b_in + b_out + b_c;
  anufft_status st = ANUFFT_OK;
  do {
    if ((e = hipMemcpyAsync(d_in, hin.data(), 8 * (size_t)nin, hipMemcpyHostToDevice, s)) != hipSuccess) break;
    if ((e = hipMemcpyAsync(d_c, hc.data(), 4 * hc.size(), hipMemcpyHostToDevice, s)) != hipSuccess) break;
    const float *dc[3] = {d_c, p->dim == 3 ? d_c + M : d_c, p->dim == 3 ? d_c + 2 * M : d_c};
    st = type == 1 ? anufft_execute_type1(p, M, dc, 1, (const float *)d_in, (float *)d_out, d_ws, wsb, s)
                   : anufft_execute_type2(p, M, dc, 1, (const float *)d_in, (float *)d_out, d_ws, wsb, s);
    if (st != ANUFFT_OK) break;
    if ((e = hipMemcpyAsync(hout.data(), d_out, 8 * (size_t)nout, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
    e = hipStreamSynchronize(s);
  } while (0);
  (void)hipStreamSynchronize(s);
  (void)hipFree(buf);
  if (st != ANUFFT_OK) return st;
  if (e != hipSuccess) return hip_fail(e, "host execute");
  from_c64(storage, hout, nout, out);
  return ANUFFT_OK;
}

// Exact direct sums from host pointers: FP64 coordinates and complex128
// data copied to the device unchanged, FP64 arithmetic, complex128 out.
anufft_status host_direct(const anufft_plan *p, int type, int64_t M, const double *const coords[3],
                          int64_t stride, const double *in, double *out) {
  if (!p) return fail(ANUFFT_INVALID_ARG, "plan is NULL");
  anufft_status sp = check_points(M, stride);
  if (sp != ANUFFT_OK) return sp;
  if (M > 0 && (!coords || !coords[0] || (p->dim == 3 && (!coords[1] || !coords[2]))))
    return fail(ANUFFT_INVALID_ARG, "coordinate pointer is NULL");
  if ((M > 0 || type == 1) && (!in || !out)) return fail(ANUFFT_INVALID_ARG, "data pointer is NULL");
  if (p->device < 0) return fail(ANUFFT_DEVICE_UNAVAILABLE, "no HIP device available for this plan");
  if (type == 2 && M == 0) return ANUFFT_OK;
  DeviceGuard guard(p->device);
  hipStream_t s = p->host_stream;
  const int64_t N = p->modes();
  const int64_t nin = type == 1 ? M : N, nout = type == 1 ? N : M;
  std::vector<double> hc((size_t)std::max<int64_t>(M, 1) * p->dim);
  for (int a = 0; a < p->dim; ++a)
    for (int64_t j = 0; j < M; ++j) hc[(size_t)a * M + j] = coords[a][j * stride];
  size_t b_in = align_up(16 * (size_t)std::max<int64_t>(nin, 1));
  size_t b_out = align_up(16 * (size_t)std::max<int64_t>(nout, 1));
  size_t b_c = align_up(8 * hc.size());
  char *buf = nullptr;
  hipError_t e = hipMalloc(&buf, b_in + b_out + b_c);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  double2 *d_in = (double2 *)buf, *d_out = (double2 *)(buf + b_in);
  double *d_c = (double *)(buf + b_in + b_out);
  do {
    if (nin > 0 && (e = hipMemcpyAsync(d_in, in, 16 * (size_t)nin, hipMemcpyHostToDevice, s)) != hipSuccess) break;
    if ((e = hipMemcpyAsync(d_c, hc.data(), 8 * hc.size(), hipMemcpyHostToDevice, s)) != hipSuccess) break;
    const double *dx = d_c, *dy = p->dim == 3 ? d_c + M : d_c, *dz = p->dim == 3 ? d_c + 2 * M : d_c;
    if ((e = launch_direct_f64(type, dx, dy, dz, 1, M, p->dim, p->n, p->L, d_in, d_out, s)) != hipSuccess) break;
    if ((e = hipMemcpyAsync(out, d_out, 16 * (size_t)nout, hipMemcpyDeviceToHost, s)) != hipSuccess) break;
    e = hipStreamSynchronize(s);
  } while (0);
  (void)hipStreamSynchronize(s);
  (void)hipFree(buf);
  if (e != hipSuccess) return hip_fail(e, "direct sum (host)");
  return ANUFFT_OK;
}

} // namespace

extern "C" {

anufft_status anufft_direct_type1_host(const anufft_plan *p, int64_t M,
                                       const double *const coords[3], int64_t stride,
                                       const double *values, double *out) {
  return host_direct(p, 1, M, coords, stride, values, out);
}

anufft_status anufft_direct_type2_host(const anufft_plan *p, int64_t M,
                                       const double *const coords[3], int64_t stride,
                                       const double *modes, double *out) {
  return host_direct(p, 2, M, coords, stride, modes, out);
}

anufft_status anufft_execute_type1_host(const anufft_plan *p, int32_t storage, int32_t profile,
                                        int64_t M, const double *const coords[3], int64_t stride,
                                        const void *values, void *out) {
  return host_exec(p, 1, storage, profile, M, coords, stride, values, out);
}

anufft_status anufft_execute_type2_host(const anufft_plan *p, int32_t storage, int32_t profile,
                                        int64_t M, const double *const coords[3], int64_t stride,
                                        const void *modes, void *out) {
  return host_exec(p, 2, storage, profile, M, coords, stride, modes, out);
}

} // extern "C"
