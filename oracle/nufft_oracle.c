/*
 * nufft_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of apollo-nufft's Kaiser-Bessel NUFFT (the reference's
 * algorithm, written here in plain C) plus FP64 direct sums.  It is the checker
 * for the HIP library and the CPU baseline timed by bench.py; it is never part
 * of the shipped device path.  Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg load it.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the apollofft repository root).  Arithmetic follows the reference exactly:
 * f64 everywhere, no FMA contraction (built with -ffp-contract=off, matching
 * Rust, which never fuses a*b+c on its own), Rust `round` = C `round`
 * (half away from zero), Rust f64 `rem_euclid` restated in or_rem_euclid().
 *
 * Complex arrays are interleaved (re, im) doubles, like num_complex::Complex64.
 * 3D arrays are row-major [x][y][z] (z fastest), like ndarray Array3 (mx,my,mz).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  double re, im;
} cpx;

static inline cpx cadd(cpx a, cpx b) { return (cpx){a.re + b.re, a.im + b.im}; }
static inline cpx cmul(cpx a, cpx b) {
  return (cpx){a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
static inline cpx cscale(cpx a, double s) { return (cpx){a.re * s, a.im * s}; }

/* ------------------------------------------------------------------------ */
/* Kaiser-Bessel primitives                                                  */
/* ------------------------------------------------------------------------ */

/* I0 power series; apollo-nufft/src/infrastructure/kernel/kaiser_bessel.rs:55-69 */
double or_i0(double z) {
  double scaled = 0.25 * z * z;
  double sum = 1.0, term = 1.0;
  for (int k = 1; k <= 256; ++k) {
    double kf = (double)k;
    term *= scaled / (kf * kf);
    sum += term;
    if (term <= 2.220446049250313e-16 * sum) break; /* f64::EPSILON */
  }
  return sum;
}

/* phi(x) = I0(beta*sqrt(1-(x/w)^2))/I0(beta); kaiser_bessel.rs:109-116 */
double or_kb_kernel(double x, double w, double beta, double i0_beta) {
  double u2 = (x / w) * (x / w);
  if (u2 >= 1.0) return 0.0;
  return or_i0(beta * sqrt(1.0 - u2)) / i0_beta;
}

/* Analytic transform of phi; kaiser_bessel.rs:120-133 */
double or_kb_kernel_ft(double xi, int64_t w, double beta, double i0_beta) {
  double two_pi_w_xi = 2.0 * M_PI * (double)w * xi;
  double z_sq = beta * beta - two_pi_w_xi * two_pi_w_xi;
  double prefix = 2.0 * (double)w / i0_beta;
  if (fabs(z_sq) < 1e-30) return prefix;
  if (z_sq > 0.0) {
    double s = sqrt(z_sq);
    return prefix * sinh(s) / s;
  }
  double s = sqrt(-z_sq);
  return prefix * sin(s) / s;
}

/* k <= len/2 -> k, else k-len; kaiser_bessel.rs:152-158 */
int64_t or_fft_signed_index(int64_t index, int64_t len) {
  return index <= len / 2 ? index : index - len;
}

/* D[k] = 1/phi_hat(k_signed/m); kaiser_bessel.rs:137-148 */
void or_axis_deconv(int64_t n, int64_t m, int64_t w, double beta, double i0_beta,
                    double *out) {
  for (int64_t k = 0; k < n; ++k) {
    double xi = (double)or_fft_signed_index(k, n) / (double)m;
    out[k] = 1.0 / or_kb_kernel_ft(xi, w, beta, i0_beta);
  }
}

/* ceil(len/(2w+1)) buckets; kaiser_bessel.rs:162-164 */
int64_t or_bucket_count(int64_t len, int64_t w) {
  int64_t l = len > 1 ? len : 1;
  int64_t d = 2 * w + 1 > 1 ? 2 * w + 1 : 1;
  int64_t b = (l + d - 1) / d;
  return b > 1 ? b : 1;
}

/* beta = pi*(1-1/(2 sigma))*2w with the integer sigma;
 * dimension_1d.rs:132, dimension_3d.rs:200 */
double or_beta(int64_t sigma, int64_t w) {
  return M_PI * (1.0 - 1.0 / (2.0 * (double)sigma)) * (double)(2 * w);
}

/* Oversampled length.
 * policy 0: 3D reference rule next_pow2(max(sigma n, 2w+1)), dimension_3d.rs:48-54
 * policy 1: exact max(sigma n, 2w+1) (mixed-radix grids such as 200, 254)
 * policy 2: 1D reference rule sigma n, dimension_1d.rs:131 */
int64_t or_oversampled_len(int64_t n, int64_t sigma, int64_t w, int policy) {
  int64_t m = n * sigma;
  if (policy == 2) return m;
  if (m < 2 * w + 1) m = 2 * w + 1;
  if (policy == 1) return m;
  int64_t p = 1;
  while (p < m) p <<= 1;
  return p;
}

static inline double or_rem_euclid(double x, double l) {
  double r = fmod(x, l); /* Rust f64 % */
  return r < 0.0 ? r + fabs(l) : r;
}
static inline int64_t or_irem(int64_t a, int64_t m) {
  int64_t r = a % m;
  return r < 0 ? r + m : r;
}

/* ------------------------------------------------------------------------ */
/* FP64 complex FFT of any length (forward sign -1, inverse sign +1,         */
/* unnormalized).  Mixed-radix decimation in time by smallest prime factor;  */
/* stands in for apollo-fft FftPlan1D<f64>                                   */
/* (apollo-fft/src/application/execution/plan/fft/dimension_1d.rs:56-82).   */
/* ------------------------------------------------------------------------ */

typedef struct {
  int64_t n;
  int sign;
  cpx *tw; /* exp(sign*2*pi*i*j/n), j in [0,n) */
  cpx *tmp;
  int64_t tmp_len;
} fft_plan;

static int64_t smallest_factor(int64_t n) {
  if (n % 4 == 0) return 4;
  if (n % 2 == 0) return 2;
  for (int64_t p = 3; p * p <= n; p += 2)
    if (n % p == 0) return p;
  return n;
}

static void fft_plan_init(fft_plan *p, int64_t n, int sign) {
  p->n = n;
  p->sign = sign;
  p->tw = (cpx *)malloc(sizeof(cpx) * (size_t)n);
  for (int64_t j = 0; j < n; ++j) {
    double a = (double)sign * 2.0 * M_PI * (double)j / (double)n;
    p->tw[j] = (cpx){cos(a), sin(a)};
  }
  p->tmp_len = n;
  p->tmp = (cpx *)malloc(sizeof(cpx) * (size_t)(n + 64));
}

static void fft_plan_free(fft_plan *p) {
  free(p->tw);
  free(p->tmp);
}

/* out[0..n) = DFT_n of in[0], in[s], ..., in[(n-1)s]; twiddle table of the
 * top-level length N, so W_n^e = tw[e*(N/n) mod N]. */
static void fft_rec(const fft_plan *pl, const cpx *in, cpx *out, int64_t n, int64_t s) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  int64_t p = smallest_factor(n);
  int64_t q = n / p;
  for (int64_t r = 0; r < p; ++r) fft_rec(pl, in + r * s, out + r * q, q, s * p);
  int64_t N = pl->n, step = N / n;
  cpx t[128];
  cpx *tt = p <= 128 ? t : (cpx *)malloc(sizeof(cpx) * (size_t)p);
  for (int64_t k = 0; k < q; ++k) {
    for (int64_t r = 0; r < p; ++r) tt[r] = cmul(out[r * q + k], pl->tw[((r * k) % n) * step]);
    if (p == 2) {
      out[k] = cadd(tt[0], tt[1]);
      out[k + q] = (cpx){tt[0].re - tt[1].re, tt[0].im - tt[1].im};
    } else if (p == 4) {
      cpx a0 = cadd(tt[0], tt[2]), a1 = (cpx){tt[0].re - tt[2].re, tt[0].im - tt[2].im};
      cpx b0 = cadd(tt[1], tt[3]), b1 = (cpx){tt[1].re - tt[3].re, tt[1].im - tt[3].im};
      /* multiply b1 by W_4^1 = sign*i */
      cpx b1r = pl->sign < 0 ? (cpx){b1.im, -b1.re} : (cpx){-b1.im, b1.re};
      out[k] = cadd(a0, b0);
      out[k + q] = cadd(a1, b1r);
      out[k + 2 * q] = (cpx){a0.re - b0.re, a0.im - b0.im};
      out[k + 3 * q] = (cpx){a1.re - b1r.re, a1.im - b1r.im};
    } else {
      int64_t pstep = N / p;
      for (int64_t u = 0; u < p; ++u) {
        cpx acc = {0.0, 0.0};
        for (int64_t r = 0; r < p; ++r) acc = cadd(acc, cmul(tt[r], pl->tw[((r * u) % p) * pstep]));
        out[k + u * q] = acc;
      }
    }
  }
  if (tt != t) free(tt);
}

static void fft_exec(fft_plan *pl, cpx *data) {
  memcpy(pl->tmp, data, sizeof(cpx) * (size_t)pl->n);
  fft_rec(pl, pl->tmp, data, pl->n, 1);
}

/* In-place unnormalized FFT of one interleaved-complex line. */
void or_fft_inplace(double *data, int64_t n, int sign) {
  fft_plan pl;
  fft_plan_init(&pl, n, sign < 0 ? -1 : 1);
  fft_exec(&pl, (cpx *)data);
  fft_plan_free(&pl);
}

/* Threads of the 3D fast restatement (1 = the reference's single-threaded
 * path).  With more, the CPU baseline's multi-thread column runs the same
 * algorithm split over points (spreading into per-thread grids, then a
 * reduction), FFT lanes and interpolated points. */
static int g_fast_threads = 1;

void or_set_fast_threads(int n) { g_fast_threads = n > 1 ? n : 1; }

/* Separable 3D FFT over a [mx][my][mz] grid, axis order given.
 * Forward: z, y, x (dimension_3d.rs:396-398, fft_{z,y,x}_pass :631-674).
 * Inverse: x, y, z (dimension_3d.rs:549-551, ifft_{x,y,z}_pass :416-459),
 * each inverse axis normalized by 1/m like FftPlan1D::inverse
 * (apollo-fft .../plan/fft/dimension_1d.rs:75-82). */
static void grid_fft3(cpx *g, const int64_t m[3], int sign, int normalize) {
  int64_t mx = m[0], my = m[1], mz = m[2];
  int64_t lmax = mx > my ? (mx > mz ? mx : mz) : (my > mz ? my : mz);
  cpx *lane = (cpx *)malloc(sizeof(cpx) * (size_t)lmax);
  int order_fwd[3] = {2, 1, 0}, order_inv[3] = {0, 1, 2};
  int *order = sign < 0 ? order_fwd : order_inv;
  for (int oi = 0; oi < 3; ++oi) {
    int ax = order[oi];
    int64_t len = m[ax];
    fft_plan pl;
    fft_plan_init(&pl, len, sign);
    double sc = normalize ? 1.0 / (double)len : 1.0;
    int64_t stride = ax == 0 ? my * mz : (ax == 1 ? mz : 1);
    int64_t na = ax == 0 ? my : mx, nb = ax == 2 ? my : mz;
    if (g_fast_threads > 1) {
#pragma omp parallel num_threads(g_fast_threads)
      {
        fft_plan tp;
        fft_plan_init(&tp, len, sign);
        cpx *tl = (cpx *)malloc(sizeof(cpx) * (size_t)lmax);
#pragma omp for schedule(static)
        for (int64_t ab = 0; ab < na * nb; ++ab) {
          int64_t a = ab / nb, b = ab % nb, base;
          if (ax == 0) base = a * mz + b;
          else if (ax == 1) base = a * my * mz + b;
          else base = a * my * mz + b * mz;
          for (int64_t i = 0; i < len; ++i) tl[i] = g[base + i * stride];
          fft_exec(&tp, tl);
          for (int64_t i = 0; i < len; ++i) g[base + i * stride] = cscale(tl[i], sc);
        }
        free(tl);
        fft_plan_free(&tp);
      }
      fft_plan_free(&pl);
      continue;
    }
    for (int64_t a = 0; a < na; ++a)
      for (int64_t b = 0; b < nb; ++b) {
        int64_t base;
        if (ax == 0) base = a * mz + b;
        else if (ax == 1) base = a * my * mz + b;
        else base = a * my * mz + b * mz;
        for (int64_t i = 0; i < len; ++i) lane[i] = g[base + i * stride];
        fft_exec(&pl, lane);
        for (int64_t i = 0; i < len; ++i) g[base + i * stride] = cscale(lane[i], sc);
      }
    fft_plan_free(&pl);
  }
  free(lane);
}

/* Plain unnormalized 3D FFT (tests use it to pin the device FFT). */
void or_fft3_inplace(double *grid, const int64_t m[3], int sign) {
  grid_fft3((cpx *)grid, m, sign < 0 ? -1 : 1, 0);
}

/* ------------------------------------------------------------------------ */
/* 1D fast NUFFT: NufftPlan1D (dimension_1d.rs:127-276)                      */
/* ------------------------------------------------------------------------ */

typedef struct {
  double x;
  cpx v;
  int64_t bucket;
  int64_t index;
} pt1;

static int cmp_pt1(const void *a, const void *b) {
  const pt1 *l = (const pt1 *)a, *r = (const pt1 *)b;
  return (l->bucket > r->bucket) - (l->bucket < r->bucket);
}

/* sort_points_1d / sort_positions_1d; dimension_1d.rs:54-99 */
static pt1 *sort_points_1d(const double *x, const double *c, int64_t M, double L,
                           int64_t m, int64_t w) {
  int64_t buckets = or_bucket_count(m, w);
  double scale = (double)buckets / L;
  pt1 *p = (pt1 *)malloc(sizeof(pt1) * (size_t)(M > 0 ? M : 1));
  for (int64_t j = 0; j < M; ++j) {
    double xm = or_rem_euclid(x[j], L);
    int64_t b = (int64_t)floor(xm * scale);
    if (b > buckets - 1) b = buckets - 1;
    p[j].x = xm;
    p[j].v = c ? (cpx){c[2 * j], c[2 * j + 1]} : (cpx){0, 0};
    p[j].bucket = b;
    p[j].index = j;
  }
  qsort(p, (size_t)M, sizeof(pt1), cmp_pt1);
  return p;
}

/* Type-1: dimension_1d.rs:166-207. out has 2n doubles. */
int or_nufft1d_type1_fast(int64_t n, double dx, int64_t sigma, int64_t w, int64_t M,
                          const double *x, const double *c, double *out) {
  if (sigma < 2 || w < 2 || n <= 0 || !(dx > 0.0)) return 1;
  int64_t m = or_oversampled_len(n, sigma, w, 2);
  double L = (double)n * dx, beta = or_beta(sigma, w), i0b = or_i0(beta);
  double *deconv = (double *)malloc(sizeof(double) * (size_t)n);
  for (int64_t k = 0; k < n; ++k) {
    double xi = (double)or_fft_signed_index(k, n) / (double)m;
    deconv[k] = 1.0 / or_kb_kernel_ft(xi, w, beta, i0b);
  }
  cpx *grid = (cpx *)calloc((size_t)m, sizeof(cpx));
  pt1 *pts = sort_points_1d(x, c, M, L, m, w);
  for (int64_t j = 0; j < M; ++j) {
    double t = (double)m * pts[j].x / L;
    int64_t m0 = (int64_t)round(t);
    double d = t - (double)m0;
    for (int64_t p = -w; p <= w; ++p) {
      double wt = or_kb_kernel((double)p - d, (double)w, beta, i0b);
      if (wt != 0.0) {
        int64_t idx = or_irem(m0 + p, m);
        grid[idx] = cadd(grid[idx], cscale(pts[j].v, wt));
      }
    }
  }
  or_fft_inplace((double *)grid, m, -1);
  cpx *o = (cpx *)out;
  for (int64_t k = 0; k < n; ++k) {
    int64_t idx = or_irem(or_fft_signed_index(k, n), m);
    o[k] = cscale(grid[idx], deconv[k]);
  }
  free(pts);
  free(grid);
  free(deconv);
  return 0;
}

/* Type-2: dimension_1d.rs:224-276 (normalized inverse FFT, then x m). */
int or_nufft1d_type2_fast(int64_t n, double dx, int64_t sigma, int64_t w, const double *f,
                          int64_t M, const double *x, double *out) {
  if (sigma < 2 || w < 2 || n <= 0 || !(dx > 0.0)) return 1;
  int64_t m = or_oversampled_len(n, sigma, w, 2);
  double L = (double)n * dx, beta = or_beta(sigma, w), i0b = or_i0(beta);
  cpx *grid = (cpx *)calloc((size_t)m, sizeof(cpx));
  for (int64_t k = 0; k < n; ++k) {
    double xi = (double)or_fft_signed_index(k, n) / (double)m;
    double dk = 1.0 / or_kb_kernel_ft(xi, w, beta, i0b);
    int64_t idx = or_irem(or_fft_signed_index(k, n), m);
    grid[idx] = cscale((cpx){f[2 * k], f[2 * k + 1]}, dk);
  }
  or_fft_inplace((double *)grid, m, +1);
  for (int64_t i = 0; i < m; ++i) grid[i] = cscale(cscale(grid[i], 1.0 / (double)m), (double)m);
  pt1 *pts = sort_points_1d(x, NULL, M, L, m, w);
  cpx *o = (cpx *)out;
  for (int64_t j = 0; j < M; ++j) {
    double t = (double)m * pts[j].x / L;
    int64_t m0 = (int64_t)round(t);
    double d = t - (double)m0;
    cpx acc = {0.0, 0.0};
    for (int64_t p = -w; p <= w; ++p) {
      double wt = or_kb_kernel((double)p - d, (double)w, beta, i0b);
      if (wt != 0.0) acc = cadd(acc, cscale(grid[or_irem(m0 + p, m)], wt));
    }
    o[pts[j].index] = acc;
  }
  free(pts);
  free(grid);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* 3D fast NUFFT: NufftPlan3D (dimension_3d.rs:193-602)                      */
/* ------------------------------------------------------------------------ */

typedef struct {
  double x, y, z;
  cpx v;
  int64_t bucket;
  int64_t index;
} pt3;

/* (bucket, x, y, z) ordering; dimension_3d.rs:109-115 */
static int cmp_pt3(const void *a, const void *b) {
  const pt3 *l = (const pt3 *)a, *r = (const pt3 *)b;
  if (l->bucket != r->bucket) return l->bucket < r->bucket ? -1 : 1;
  if (l->x != r->x) return l->x < r->x ? -1 : 1;
  if (l->y != r->y) return l->y < r->y ? -1 : 1;
  if (l->z != r->z) return l->z < r->z ? -1 : 1;
  return 0;
}

/* sort_points_3d / sort_positions_3d; dimension_3d.rs:74-161 */
static pt3 *sort_points_3d(const double *xyz, const double *c, int64_t M, const double L[3],
                           const int64_t m[3], int64_t w) {
  int64_t bx = or_bucket_count(m[0], w), by = or_bucket_count(m[1], w),
          bz = or_bucket_count(m[2], w);
  double sx = (double)bx / L[0], sy = (double)by / L[1], sz = (double)bz / L[2];
  pt3 *p = (pt3 *)malloc(sizeof(pt3) * (size_t)(M > 0 ? M : 1));
  for (int64_t j = 0; j < M; ++j) {
    double xm = or_rem_euclid(xyz[3 * j], L[0]);
    double ym = or_rem_euclid(xyz[3 * j + 1], L[1]);
    double zm = or_rem_euclid(xyz[3 * j + 2], L[2]);
    int64_t ix = (int64_t)floor(xm * sx), iy = (int64_t)floor(ym * sy),
            iz = (int64_t)floor(zm * sz);
    if (ix > bx - 1) ix = bx - 1;
    if (iy > by - 1) iy = by - 1;
    if (iz > bz - 1) iz = bz - 1;
    p[j] = (pt3){xm, ym, zm, c ? (cpx){c[2 * j], c[2 * j + 1]} : (cpx){0, 0},
                 (ix * by + iy) * bz + iz, j};
  }
  qsort(p, (size_t)M, sizeof(pt3), cmp_pt3);
  return p;
}

typedef struct {
  int64_t n[3], m[3], w;
  double L[3], beta, i0b;
  double *deconv[3];
} plan3;

static int plan3_init(plan3 *pl, const int64_t n[3], const double d[3], int64_t sigma,
                      int64_t w, int policy) {
  if (sigma < 2 || w < 2) return 1;
  for (int a = 0; a < 3; ++a)
    if (n[a] <= 0 || !(d[a] > 0.0)) return 1;
  pl->w = w;
  pl->beta = or_beta(sigma, w);
  pl->i0b = or_i0(pl->beta);
  for (int a = 0; a < 3; ++a) {
    pl->n[a] = n[a];
    pl->m[a] = or_oversampled_len(n[a], sigma, w, policy);
    pl->L[a] = (double)n[a] * d[a];
    pl->deconv[a] = (double *)malloc(sizeof(double) * (size_t)n[a]);
    or_axis_deconv(n[a], pl->m[a], w, pl->beta, pl->i0b, pl->deconv[a]);
  }
  return 0;
}

static void plan3_free(plan3 *pl) {
  for (int a = 0; a < 3; ++a) free(pl->deconv[a]);
}

/* Per-point tap weights: (2w+1) kb_kernel calls per axis, dimension_3d.rs:291-305 */
static void point_weights(const plan3 *pl, const pt3 *p, int64_t m0[3], double *wx,
                          double *wy, double *wz) {
  double t[3] = {(double)pl->m[0] * p->x / pl->L[0], (double)pl->m[1] * p->y / pl->L[1],
                 (double)pl->m[2] * p->z / pl->L[2]};
  double d[3];
  for (int a = 0; a < 3; ++a) {
    m0[a] = (int64_t)round(t[a]);
    d[a] = t[a] - (double)m0[a];
  }
  int64_t w = pl->w;
  for (int64_t p_ = -w, o = 0; p_ <= w; ++p_, ++o) {
    wx[o] = or_kb_kernel((double)p_ - d[0], (double)w, pl->beta, pl->i0b);
    wy[o] = or_kb_kernel((double)p_ - d[1], (double)w, pl->beta, pl->i0b);
    wz[o] = or_kb_kernel((double)p_ - d[2], (double)w, pl->beta, pl->i0b);
  }
}

/* The T1 hot loop over points (dimension_3d.rs:290-327): per point the
 * (2w+1)^3 taps with zero-skip per axis, added into grid. */
static void spread_points(const plan3 *pl, const pt3 *pts, int64_t M, cpx *grid) {
  const int64_t mx = pl->m[0], my = pl->m[1], mz = pl->m[2], w = pl->w;
  double wx[64], wy[64], wz[64];
  int64_t m0[3];
  for (int64_t j = 0; j < M; ++j) {
    point_weights(pl, &pts[j], m0, wx, wy, wz);
    for (int64_t px = 0; px <= 2 * w; ++px) {
      if (wx[px] == 0.0) continue;
      int64_t ix = or_irem(m0[0] + px - w, mx);
      for (int64_t py = 0; py <= 2 * w; ++py) {
        if (wy[py] == 0.0) continue;
        int64_t iy = or_irem(m0[1] + py - w, my);
        double wxy = wx[px] * wy[py];
        cpx *row = grid + (ix * my + iy) * mz;
        for (int64_t pz = 0; pz <= 2 * w; ++pz) {
          if (wz[pz] == 0.0) continue;
          int64_t iz = or_irem(m0[2] + pz - w, mz);
          row[iz] = cadd(row[iz], cscale(pts[j].v, wxy * wz[pz]));
        }
      }
    }
  }
}

/* Spread onto the oversampled grid (T1 hot loop, dimension_3d.rs:277-327).
 * Exposed so tests can pin the device spreader on the fine grid itself. */
int or_spread3d(const int64_t n[3], const double d[3], int64_t sigma, int64_t w, int policy,
                int64_t M, const double *xyz, const double *c, double *grid_out) {
  plan3 pl;
  if (plan3_init(&pl, n, d, sigma, w, policy)) return 1;
  int64_t mx = pl.m[0], my = pl.m[1], mz = pl.m[2];
  cpx *grid = (cpx *)grid_out;
  memset(grid, 0, sizeof(cpx) * (size_t)(mx * my * mz));
  pt3 *pts = sort_points_3d(xyz, c, M, pl.L, pl.m, w);
  if (g_fast_threads > 1) {
    /* contiguous runs of the sorted points into per-thread grids, then a
     * reduction over cells (multi-thread CPU baseline only) */
    const int64_t G = mx * my * mz;
    const int T = g_fast_threads;
    cpx **priv = (cpx **)malloc(sizeof(cpx *) * (size_t)T);
#pragma omp parallel num_threads(T)
    {
      const int t = omp_get_thread_num(), nt = omp_get_num_threads();
      cpx *mine = t == 0 ? grid : (cpx *)calloc((size_t)G, sizeof(cpx));
      priv[t] = mine;
      const int64_t j0 = M * t / nt, j1 = M * (t + 1) / nt;
      spread_points(&pl, pts + j0, j1 - j0, mine);
#pragma omp barrier
#pragma omp for schedule(static)
      for (int64_t i = 0; i < G; ++i)
        for (int u = 1; u < nt; ++u) grid[i] = cadd(grid[i], priv[u][i]);
#pragma omp barrier
      if (t > 0) free(mine);
    }
    free(priv);
  } else {
    spread_points(&pl, pts, M, grid);
  }
  free(pts);
  plan3_free(&pl);
  return 0;
}

/* Type-1 3D: dimension_3d.rs:256-330 and forward_oversampled_grid_into :391-414. */
int or_nufft3d_type1_fast(const int64_t n[3], const double d[3], int64_t sigma, int64_t w,
                          int policy, int64_t M, const double *xyz, const double *c,
                          double *out) {
  plan3 pl;
  if (plan3_init(&pl, n, d, sigma, w, policy)) return 1;
  int64_t mx = pl.m[0], my = pl.m[1], mz = pl.m[2];
  cpx *grid = (cpx *)malloc(sizeof(cpx) * (size_t)(mx * my * mz));
  or_spread3d(n, d, sigma, w, policy, M, xyz, c, (double *)grid);
  grid_fft3(grid, pl.m, -1, 0);
  cpx *o = (cpx *)out;
  for (int64_t kx = 0; kx < n[0]; ++kx)
    for (int64_t ky = 0; ky < n[1]; ++ky)
      for (int64_t kz = 0; kz < n[2]; ++kz) {
        int64_t ix = or_irem(or_fft_signed_index(kx, n[0]), mx);
        int64_t iy = or_irem(or_fft_signed_index(ky, n[1]), my);
        int64_t iz = or_irem(or_fft_signed_index(kz, n[2]), mz);
        double dk = pl.deconv[0][kx] * pl.deconv[1][ky] * pl.deconv[2][kz];
        o[(kx * n[1] + ky) * n[2] + kz] = cscale(grid[(ix * my + iy) * mz + iz], dk);
      }
  free(grid);
  plan3_free(&pl);
  return 0;
}

/* Type-2 3D: dimension_3d.rs:496-602.  compat_scaling != 0 reproduces the
 * reference output exactly (three normalized inverse passes, no x G rescale,
 * so the result is the direct sum / (mx my mz)); compat_scaling == 0 rescales
 * by mx my mz so the result equals the direct sum. */
int or_nufft3d_type2_fast(const int64_t n[3], const double d[3], int64_t sigma, int64_t w,
                          int policy, int compat_scaling, const double *f, int64_t M,
                          const double *xyz, double *out) {
  plan3 pl;
  if (plan3_init(&pl, n, d, sigma, w, policy)) return 1;
  int64_t mx = pl.m[0], my = pl.m[1], mz = pl.m[2];
  cpx *grid = (cpx *)calloc((size_t)(mx * my * mz), sizeof(cpx));
  for (int64_t kx = 0; kx < n[0]; ++kx) {
    int64_t ix = or_irem(or_fft_signed_index(kx, n[0]), mx);
    for (int64_t ky = 0; ky < n[1]; ++ky) {
      int64_t iy = or_irem(or_fft_signed_index(ky, n[1]), my);
      for (int64_t kz = 0; kz < n[2]; ++kz) {
        int64_t iz = or_irem(or_fft_signed_index(kz, n[2]), mz);
        int64_t k = (kx * n[1] + ky) * n[2] + kz;
        double dk = pl.deconv[0][kx] * pl.deconv[1][ky] * pl.deconv[2][kz];
        grid[(ix * my + iy) * mz + iz] = cscale((cpx){f[2 * k], f[2 * k + 1]}, dk);
      }
    }
  }
  grid_fft3(grid, pl.m, +1, 1);
  if (!compat_scaling) {
    double g = (double)mx * (double)my * (double)mz;
    for (int64_t i = 0; i < mx * my * mz; ++i) grid[i] = cscale(grid[i], g);
  }
  pt3 *pts = sort_points_3d(xyz, NULL, M, pl.L, pl.m, w);
  cpx *o = (cpx *)out;
#pragma omp parallel for schedule(static) num_threads(g_fast_threads) if (g_fast_threads > 1)
  for (int64_t j = 0; j < M; ++j) {
    double wx[64], wy[64], wz[64];
    int64_t m0[3];
    point_weights(&pl, &pts[j], m0, wx, wy, wz);
    cpx acc = {0.0, 0.0};
    for (int64_t px = 0; px <= 2 * w; ++px) {
      if (wx[px] == 0.0) continue;
      int64_t ix = or_irem(m0[0] + px - w, mx);
      for (int64_t py = 0; py <= 2 * w; ++py) {
        if (wy[py] == 0.0) continue;
        int64_t iy = or_irem(m0[1] + py - w, my);
        double wxy = wx[px] * wy[py];
        const cpx *row = grid + (ix * my + iy) * mz;
        for (int64_t pz = 0; pz <= 2 * w; ++pz) {
          if (wz[pz] == 0.0) continue;
          acc = cadd(acc, cscale(row[or_irem(m0[2] + pz - w, mz)], wxy * wz[pz]));
        }
      }
    }
    o[pts[j].index] = acc;
  }
  free(pts);
  free(grid);
  plan3_free(&pl);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* FP64 direct sums: the definition every fast result is measured against.  */
/* ------------------------------------------------------------------------ */

/* f_k = sum_j c_j exp(-2 pi i k_signed x_j / L); dimension_1d.rs:429-450 */
void or_direct_type1_1d(int64_t n, double dx, int64_t M, const double *x, const double *c,
                        double *out) {
  double L = (double)n * dx, two_pi_over_l = 2.0 * M_PI / L;
  cpx *o = (cpx *)out;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < n; ++k) {
    double ks = (double)or_fft_signed_index(k, n);
    cpx acc = {0.0, 0.0};
    for (int64_t j = 0; j < M; ++j) {
      double a = -two_pi_over_l * ks * x[j];
      acc = cadd(acc, cmul((cpx){c[2 * j], c[2 * j + 1]}, (cpx){cos(a), sin(a)}));
    }
    o[k] = acc;
  }
}

/* g_j = sum_k f_k exp(+2 pi i k_signed x_j / L); dimension_1d.rs:454-472 */
void or_direct_type2_1d(int64_t n, double dx, const double *f, int64_t M, const double *x,
                        double *out) {
  double L = (double)n * dx;
  cpx *o = (cpx *)out;
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < M; ++j) {
    cpx acc = {0.0, 0.0};
    for (int64_t k = 0; k < n; ++k) {
      double a = 2.0 * M_PI * (double)or_fft_signed_index(k, n) * x[j] / L;
      acc = cadd(acc, cmul((cpx){f[2 * k], f[2 * k + 1]}, (cpx){cos(a), sin(a)}));
    }
    o[j] = acc;
  }
}

/* Powers e^{i s k theta} for k_signed in [-(n-1)/2.., n/2], stored by FFT index. */
static void axis_phases(double theta, int64_t n, cpx *tab) {
  /* tab[k] = exp(i * k_signed(k) * theta).  The non-negative powers s =
   * 0..n/2 come from a direct sincos of the exact multiple every 8 steps and
   * one complex multiplication by e^{i theta} in between (at most 7 rounding
   * steps, ~1e-15 relative); negative indices take the conjugate of the
   * matching positive power (theta is real). */
  const cpx step = {cos(theta), sin(theta)};
  const int64_t half = n / 2;
  cpx cur = {1.0, 0.0};
  for (int64_t s = 0; s <= half; ++s) {
    if ((s & 7) == 0) {
      const double a = (double)s * theta;
      cur = (cpx){cos(a), sin(a)};
    } else {
      cur = cmul(cur, step);
    }
    tab[s] = cur;
  }
  for (int64_t k = half + 1; k < n; ++k) {
    const cpx p = tab[n - k];
    tab[k] = (cpx){p.re, -p.im};
  }
}

/* Type-1 3D direct sum at selected modes (flat FFT-order indices into
 * [nx][ny][nz]); dimension_3d.rs:679-708.  Parallel over points. */
void or_direct_type1_3d(const int64_t n[3], const double d[3], int64_t M, const double *xyz,
                        const double *c, int64_t nsel, const int64_t *sel, double *out) {
  double L[3] = {(double)n[0] * d[0], (double)n[1] * d[1], (double)n[2] * d[2]};
  int64_t *kk = (int64_t *)malloc(sizeof(int64_t) * (size_t)(3 * (nsel > 0 ? nsel : 1)));
  for (int64_t s = 0; s < nsel; ++s) {
    int64_t f = sel[s];
    kk[3 * s + 2] = f % n[2];
    kk[3 * s + 1] = (f / n[2]) % n[1];
    kk[3 * s + 0] = f / (n[1] * n[2]);
  }
  cpx *o = (cpx *)out;
  for (int64_t s = 0; s < nsel; ++s) o[s] = (cpx){0.0, 0.0};
#pragma omp parallel
  {
    cpx *acc = (cpx *)calloc((size_t)(nsel > 0 ? nsel : 1), sizeof(cpx));
    cpx *ex = (cpx *)malloc(sizeof(cpx) * (size_t)n[0]);
    cpx *ey = (cpx *)malloc(sizeof(cpx) * (size_t)n[1]);
    cpx *ez = (cpx *)malloc(sizeof(cpx) * (size_t)n[2]);
#pragma omp for schedule(static)
    for (int64_t j = 0; j < M; ++j) {
      axis_phases(-2.0 * M_PI / L[0] * xyz[3 * j], n[0], ex);
      axis_phases(-2.0 * M_PI / L[1] * xyz[3 * j + 1], n[1], ey);
      axis_phases(-2.0 * M_PI / L[2] * xyz[3 * j + 2], n[2], ez);
      cpx cj = {c[2 * j], c[2 * j + 1]};
      for (int64_t s = 0; s < nsel; ++s) {
        cpx e = cmul(cmul(ex[kk[3 * s]], ey[kk[3 * s + 1]]), ez[kk[3 * s + 2]]);
        acc[s] = cadd(acc[s], cmul(cj, e));
      }
    }
#pragma omp critical
    for (int64_t s = 0; s < nsel; ++s) o[s] = cadd(o[s], acc[s]);
    free(acc);
    free(ex);
    free(ey);
    free(ez);
  }
  free(kk);
}

/* Type-2 3D direct sum at the given points; dimension_3d.rs:712-740.
 * Separable nesting: sum_kx ex sum_ky ey sum_kz ez f. */
void or_direct_type2_3d(const int64_t n[3], const double d[3], const double *f, int64_t npts,
                        const double *xyz, double *out) {
  double L[3] = {(double)n[0] * d[0], (double)n[1] * d[1], (double)n[2] * d[2]};
  const cpx *F = (const cpx *)f;
  cpx *o = (cpx *)out;
#pragma omp parallel
  {
    cpx *ex = (cpx *)malloc(sizeof(cpx) * (size_t)n[0]);
    cpx *ey = (cpx *)malloc(sizeof(cpx) * (size_t)n[1]);
    cpx *ez = (cpx *)malloc(sizeof(cpx) * (size_t)n[2]);
#pragma omp for schedule(dynamic, 4)
    for (int64_t j = 0; j < npts; ++j) {
      axis_phases(2.0 * M_PI / L[0] * xyz[3 * j], n[0], ex);
      axis_phases(2.0 * M_PI / L[1] * xyz[3 * j + 1], n[1], ey);
      axis_phases(2.0 * M_PI / L[2] * xyz[3 * j + 2], n[2], ez);
      cpx sx = {0.0, 0.0};
      for (int64_t kx = 0; kx < n[0]; ++kx) {
        cpx sy = {0.0, 0.0};
        for (int64_t ky = 0; ky < n[1]; ++ky) {
          const cpx *row = F + (kx * n[1] + ky) * n[2];
          cpx sz = {0.0, 0.0};
          for (int64_t kz = 0; kz < n[2]; ++kz) sz = cadd(sz, cmul(row[kz], ez[kz]));
          sy = cadd(sy, cmul(sz, ey[ky]));
        }
        sx = cadd(sx, cmul(sy, ex[kx]));
      }
      o[j] = sx;
    }
    free(ex);
    free(ey);
    free(ez);
  }
}

/* Direct-sum gradient of L = ||T2(f,x) - b||^2 with respect to the point
 * coordinates, at selected points (closed form, no reference counterpart):
 * dL/dx_j^(a) = 2 Re( conj(r_j) * sum_k i (2 pi/L_a) k_a f_k e^{i k.x_j} ),
 * with r = T2(f,x) - b supplied by the caller.  out is npts x 3. */
void or_direct_type2_3d_grad_x(const int64_t n[3], const double d[3], const double *f,
                               int64_t npts, const double *xyz, const double *r,
                               double *out) {
  double L[3] = {(double)n[0] * d[0], (double)n[1] * d[1], (double)n[2] * d[2]};
  const cpx *F = (const cpx *)f;
#pragma omp parallel
  {
    cpx *ex = (cpx *)malloc(sizeof(cpx) * (size_t)n[0]);
    cpx *ey = (cpx *)malloc(sizeof(cpx) * (size_t)n[1]);
    cpx *ez = (cpx *)malloc(sizeof(cpx) * (size_t)n[2]);
#pragma omp for schedule(dynamic, 4)
    for (int64_t j = 0; j < npts; ++j) {
      axis_phases(2.0 * M_PI / L[0] * xyz[3 * j], n[0], ex);
      axis_phases(2.0 * M_PI / L[1] * xyz[3 * j + 1], n[1], ey);
      axis_phases(2.0 * M_PI / L[2] * xyz[3 * j + 2], n[2], ez);
      cpx g[3] = {{0, 0}, {0, 0}, {0, 0}};
      for (int64_t kx = 0; kx < n[0]; ++kx) {
        double kxs = (double)or_fft_signed_index(kx, n[0]) * 2.0 * M_PI / L[0];
        for (int64_t ky = 0; ky < n[1]; ++ky) {
          double kys = (double)or_fft_signed_index(ky, n[1]) * 2.0 * M_PI / L[1];
          const cpx *row = F + (kx * n[1] + ky) * n[2];
          cpx exy = cmul(ex[kx], ey[ky]);
          for (int64_t kz = 0; kz < n[2]; ++kz) {
            double kzs = (double)or_fft_signed_index(kz, n[2]) * 2.0 * M_PI / L[2];
            cpx t = cmul(row[kz], cmul(exy, ez[kz]));
            g[0] = cadd(g[0], cscale(t, kxs));
            g[1] = cadd(g[1], cscale(t, kys));
            g[2] = cadd(g[2], cscale(t, kzs));
          }
        }
      }
      cpx rj = {r[2 * j], r[2 * j + 1]};
      for (int a = 0; a < 3; ++a) {
        /* i*g */
        cpx ig = {-g[a].im, g[a].re};
        /* 2 Re(conj(r) * ig) */
        out[3 * j + a] = 2.0 * (rj.re * ig.re + rj.im * ig.im);
      }
    }
    free(ex);
    free(ey);
    free(ez);
  }
}

int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
