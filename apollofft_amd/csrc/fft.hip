// fft.hip -- batched 1D FFT passes over the lines of the oversampled grid.
//
// Replaces the per-axis lane passes of apollo-nufft (fft_{x,y,z}_pass and
// ifft_{x,y,z}_pass, transform/dimension_3d.rs:416-459,631-674) and the
// WGPU GpuFft3d pack/bitrev/stage/unpack passes (apollo-fft-wgpu
// .../gpu_fft/dispatch.rs:7-215).  One launch per axis: a block loads
// `lb` lines into LDS (coalesced along whichever of the line or element
// index is contiguous in HBM), runs a mixed-radix Stockham FFT in LDS
// (radix 8/4/2 butterflies, radix 3/5 and any other prime by direct
// small DFTs), and writes the lines back.  The NUFFT's zero padding and
// cropping are fused into the load/store: a pass may read only the kept
// modes of a line (zero-filling the rest in LDS) or write only the kept
// modes, and the per-axis KB deconvolution factor (FP64 table) is applied
// on the fly, so crop/pad/deconvolve never make a separate HBM pass.
// Twiddles are FP64-built and rounded once to FP32.  Lengths with a prime
// factor above 16 use Bluestein's algorithm: the DFT of length L is a
// circular convolution of length Lf = 2^k >= 2L - 1 with the chirp
// exp(-i pi n^2 / L), done as FFT, pointwise product with the precomputed
// filter spectrum, FFT of the conjugate, all in the same LDS lines.
#include "internal.h"

namespace anufft {
namespace {

constexpr int kFftThreads = 256;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }

// x / d for x d < 2^32, with m = ceil(2^32 / d) (d > 1): floor(x m / 2^32)
// exceeds x / d by less than x / 2^32 < 1 / d, which cannot reach the next
// integer.  The index arithmetic of a pass stays below 2^13 x 2^13.
__device__ __forceinline__ int fdiv(int x, int d, uint32_t m) {
  return d == 1 ? x : (int)__umulhi((uint32_t)x, m);
}

template <int R>
__device__ __forceinline__ void small_dft(float2 *v, const float2 *tw, int L);

template <>
__device__ __forceinline__ void small_dft<2>(float2 *v, const float2 *, int) {
  float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <>
__device__ __forceinline__ void small_dft<4>(float2 *v, const float2 *, int) {
  float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
  float2 b0 = cadd(v[1], v[3]), b1 = mul_mi(csub(v[1], v[3]));
  v[0] = cadd(a0, b0);
  v[1] = cadd(a1, b1);
  v[2] = csub(a0, b0);
  v[3] = csub(a1, b1);
}

template <>
__device__ __forceinline__ void small_dft<8>(float2 *v, const float2 *, int) {
  const float h = 0.70710678118654752440f;
  // stage 1: radix-2 across the halves
  float2 a[4], b[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    a[r] = cadd(v[r], v[r + 4]);
    b[r] = csub(v[r], v[r + 4]);
  }
  // twiddles W8^r on b: 1, (1-i)h, -i, (-1-i)h
  b[1] = make_float2((b[1].x + b[1].y) * h, (b[1].y - b[1].x) * h);
  b[2] = mul_mi(b[2]);
  b[3] = make_float2((b[3].y - b[3].x) * h, -(b[3].x + b[3].y) * h);
  small_dft<4>(a, nullptr, 0);
  small_dft<4>(b, nullptr, 0);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    v[2 * r] = a[r];
    v[2 * r + 1] = b[r];
  }
}

// Radix 3 and 5 by direct sums with table roots W_R^e = tw[(e mod R) L/R].
template <int R>
__device__ __forceinline__ void direct_dft(float2 *v, const float2 *tw, int L) {
  float2 o[R];
  const int st = L / R;
#pragma unroll
  for (int u = 0; u < R; ++u) {
    float2 acc = v[0];
#pragma unroll
    for (int r = 1; r < R; ++r) acc = cadd(acc, cmul(v[r], tw[((r * u) % R) * st]));
    o[u] = acc;
  }
#pragma unroll
  for (int u = 0; u < R; ++u) v[u] = o[u];
}
template <>
__device__ __forceinline__ void small_dft<3>(float2 *v, const float2 *tw, int L) {
  direct_dft<3>(v, tw, L);
}
template <>
__device__ __forceinline__ void small_dft<5>(float2 *v, const float2 *tw, int L) {
  direct_dft<5>(v, tw, L);
}

// Stockham radix-R stage (decimation in time, autosorting):
//   j in [0, L/R), k = j mod Ns, v_r = src[j + r L/R] * W_{Ns R}^{k r},
//   dst[(j - k) R + k + u Ns] = DFT_R(v)_u.
template <int R>
__device__ __forceinline__ void stage(const float2 *src, float2 *dst, int L, int Ns,
                                      const float2 *tw, int lines, uint32_t mag_nbf,
                                      uint32_t mag_ns, int twstep) {
  const int nbf = L / R;
  for (int it = threadIdx.x; it < lines * nbf; it += kFftThreads) {
    const int li = fdiv(it, nbf, mag_nbf);
    const int j = it - li * nbf;
    const int k = j - fdiv(j, Ns, mag_ns) * Ns;
    const float2 *s = src + li * L;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[j + r * nbf];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[k * r * twstep]);
    }
    small_dft<R>(v, tw, L);
    float2 *d = dst + li * L + (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[r * Ns] = v[r];
  }
}

// Any radix (primes above 5): one thread per output element.  With
// W_R = W_L^(L/R), the twiddle of input r is W_L^(r (k twstep + u L/R)), so
// its exponent advances by a fixed step modulo L.
__device__ void stage_generic(const float2 *src, float2 *dst, int L, int Ns, int R,
                              const float2 *tw, int lines) {
  const int nbf = L / R;
  const int twstep = L / (Ns * R);
  for (int it = threadIdx.x; it < lines * L; it += kFftThreads) {
    int li = it / L;
    int o = it - li * L;
    int k = o % Ns;
    int u = (o / Ns) % R;
    int jq = o / (Ns * R);
    int j = jq * Ns + k;
    const float2 *s = src + li * L + j;
    const int step = (k * twstep + u * nbf) % L;
    float2 acc = s[0];
    int e = 0;
    for (int r = 1; r < R; ++r) {
      e += step;
      if (e >= L) e -= L;
      acc = cadd(acc, cmul(s[r * nbf], tw[e]));
    }
    dst[li * L + o] = acc;
  }
}

__device__ __forceinline__ int kept_pos(int i, int nk, int L) {
  // index i of the nk kept modes (FFT order) -> position on a length-L line
  return (nk == L || i <= nk / 2) ? i : L - nk + i;
}

// All Stockham stages of one length-Lf transform of nl lines, ping-ponging
// between *src and *dst (the result ends in *src).
__device__ __forceinline__ void run_stages(const FftPass &p, int Lf, const float2 *tw, int nl,
                                           float2 *&src, float2 *&dst) {
  int Ns = 1;
  for (int f = 0; f < p.nfac; ++f) {
    __syncthreads();
    const int R = p.fac[f];
    const uint32_t mb = p.mag_nbf[f], ms = p.mag_ns[f];
    const int ts = p.twstep[f];
    switch (R) {
    case 2: stage<2>(src, dst, Lf, Ns, tw, nl, mb, ms, ts); break;
    case 3: stage<3>(src, dst, Lf, Ns, tw, nl, mb, ms, ts); break;
    case 4: stage<4>(src, dst, Lf, Ns, tw, nl, mb, ms, ts); break;
    case 5: stage<5>(src, dst, Lf, Ns, tw, nl, mb, ms, ts); break;
    case 8: stage<8>(src, dst, Lf, Ns, tw, nl, mb, ms, ts); break;
    default: stage_generic(src, dst, Lf, Ns, R, tw, nl); break;
    }
    float2 *t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
}

// Elements of a pass handled per thread at a time: their global loads (the
// input, its deconvolution factor) are all issued before any is used.  Past
// the end a lane loads element 0 of its block and drops it, so no branch
// around a load makes the compiler wait for it there.
constexpr int kLoadBatch = 4;

// (line, index) of element idx of a block: index-major when the lines are
// strided (adjacent lines adjacent in memory), line-major otherwise.
__device__ __forceinline__ int2 line_index(int idx, bool contig, int n, uint32_t mag_n, int nl,
                                           uint32_t mag_nl) {
  if (contig) {
    const int li = fdiv(idx, n, mag_n);
    return make_int2(li, idx - li * n);
  }
  const int i = fdiv(idx, nl, mag_nl);
  return make_int2(idx - i * nl, i);
}

// Input lines into LDS: deconvolution (DEC), conjugation for the inverse,
// the Bluestein chirp, kept modes at their FFT positions.
template <bool DEC>
__device__ __forceinline__ void load_lines(const FftPass &p, const int64_t *lbase, int nl,
                                           uint32_t mag_nl, int Lf, bool blue, float2 *buf0) {
  const bool contig = p.in_si == 1;
  const int n = nl * p.nin;
  for (int base = threadIdx.x; base < n; base += kFftThreads * kLoadBatch) {
    float2 v[kLoadBatch];
    double f[kLoadBatch];
    int2 e[kLoadBatch];  // (line, index)
#pragma unroll
    for (int r = 0; r < kLoadBatch; ++r) {
      const int idx = base + r * kFftThreads;
      e[r] = line_index(idx < n ? idx : 0, contig, p.nin, p.mag_nin, nl, mag_nl);
      v[r] = p.in[lbase[e[r].x] + (int64_t)e[r].y * p.in_si];
      f[r] = DEC ? p.dec_in[e[r].y] : 1.0;
    }
#pragma unroll
    for (int r = 0; r < kLoadBatch; ++r) {
      float2 x = v[r];
      if (DEC) {
        const float s = (float)f[r];
        x.x *= s;
        x.y *= s;
      }
      if (p.inverse) x.y = -x.y;
      const int pos = kept_pos(e[r].y, p.nin, p.L);
      if (blue) x = cmul(x, p.chirp[pos]);
      if (base + r * kFftThreads < n) buf0[e[r].x * Lf + pos] = x;
    }
  }
}

// LDS lines to the output: kept modes, Bluestein chirp, conjugation for the
// inverse, scale and deconvolution (DEC), and the bin maxima (bm != null).
template <bool DEC>
__device__ __forceinline__ void store_lines(const FftPass &p, const int64_t *lbase, int nl,
                                            uint32_t mag_nl, int Lf, bool blue, const float2 *src,
                                            uint32_t *bm, int nbz) {
  const bool contig = p.out_si == 1;
  const int n = nl * p.nout;
  for (int base = threadIdx.x; base < n; base += kFftThreads * kLoadBatch) {
    double f[kLoadBatch];
    int2 e[kLoadBatch];  // (line, output index)
#pragma unroll
    for (int r = 0; r < kLoadBatch; ++r) {
      const int idx = base + r * kFftThreads;
      e[r] = line_index(idx < n ? idx : 0, contig, p.nout, p.mag_nout, nl, mag_nl);
      f[r] = DEC ? p.dec_out[e[r].y] : 1.0;
    }
#pragma unroll
    for (int r = 0; r < kLoadBatch; ++r) {
      if (base + r * kFftThreads >= n) continue;
      const int li = e[r].x, o = e[r].y;
      const int pos = kept_pos(o, p.nout, p.L);
      float2 v = src[li * Lf + pos];
      if (blue) v = cmul(make_float2(v.x, -v.y), p.chirp[pos]);  // X_k = chirp_k conj(c_k)
      if (p.inverse) v.y = -v.y;
      const float s = (float)(DEC ? p.scale * f[r] : p.scale);
      v.x *= s;
      v.y *= s;
      p.out[lbase[p.lb + li] + (int64_t)o * p.out_si] = v;
      if (bm)
        atomicMax(&bm[li * nbz + fdiv(o, p.bm_bin[2], p.mag_bmz)],
                  __float_as_uint(fmaxf(fabsf(v.x), fabsf(v.y))));
    }
  }
}

__global__ __launch_bounds__(kFftThreads) void fft_pass_kernel(FftPass p) {
  extern __shared__ float2 sh[];
  const int L = p.L, lb = p.lb;
  const bool blue = p.blue_n > 0;
  const int Lf = blue ? p.blue_n : L;  // transform length and line stride in LDS
  float2 *buf0 = sh;
  float2 *buf1 = sh + (size_t)lb * Lf;
  float2 *tw = sh + 2 * (size_t)lb * Lf;  // twiddle table of length Lf
  int64_t *lbase = (int64_t *)(tw + Lf);  // [lb] input, [lb] output line offsets
  // the first 2 x kFftThreads twiddles: loaded now, stored to LDS after the
  // input lines (their latency overlaps the line set-up and loads)
  const float2 tw0 = p.tw[min((int)threadIdx.x, Lf - 1)];
  const float2 tw1 = p.tw[min((int)threadIdx.x + kFftThreads, Lf - 1)];
  if (p.zero)
    for (int i = blockIdx.x * kFftThreads + threadIdx.x; i < p.zero_n; i += gridDim.x * kFftThreads)
      p.zero[i] = 0.f;
  const int64_t l0 = (int64_t)blockIdx.x * lb;
  const int nl = (int)((p.nlines - l0) < lb ? (p.nlines - l0) : lb);
  const uint32_t mag_nl = nl == lb ? p.mag_lb : p.mag_last;  // nl < lb only in the last block
  // element offsets of the block's lines, one 64-bit division per line
  for (int li = threadIdx.x; li < nl; li += kFftThreads) {
    const int64_t l = l0 + li;
    const int64_t a = p.use_mag_nb ? (int64_t)fdiv((int)l, (int)p.nb, p.mag_nb) : l / p.nb;
    const int64_t b = l - a * p.nb;
    lbase[li] = a * p.in_sa + b * p.in_sb;
    lbase[lb + li] = a * p.out_sa + b * p.out_sb;
  }
  __syncthreads();

  if (blue) {
    // a_n = x_n chirp_n on [0, L), zero on [L, Lf): clear everything first
    for (int idx = threadIdx.x; idx < nl * Lf; idx += kFftThreads) buf0[idx] = make_float2(0.f, 0.f);
    __syncthreads();
  }
  // ---- load (with deconvolution, conjugation for the inverse, zero pad) ----
  if (p.dec_in) load_lines<true>(p, lbase, nl, mag_nl, Lf, blue, buf0);
  else load_lines<false>(p, lbase, nl, mag_nl, Lf, blue, buf0);
  if ((int)threadIdx.x < Lf) tw[threadIdx.x] = tw0;
  if ((int)threadIdx.x + kFftThreads < Lf) tw[threadIdx.x + kFftThreads] = tw1;
  for (int i = threadIdx.x + 2 * kFftThreads; i < Lf; i += kFftThreads) tw[i] = p.tw[i];
  if (!blue && p.nin < L) {
    const int gap = L - p.nin, g0 = p.nin / 2 + 1;
    for (int idx = threadIdx.x; idx < nl * gap; idx += kFftThreads) {
      int li = idx / gap;
      buf0[li * L + g0 + (idx - li * gap)] = make_float2(0.f, 0.f);
    }
  }

  // ---- Stockham stages in LDS ----
  float2 *src = buf0, *dst = buf1;
  run_stages(p, Lf, tw, nl, src, dst);
  if (blue) {
    // C = FFT(a) FFT(b) / Lf; the inverse FFT of C is conj(FFT(conj(C)))
    __syncthreads();
    for (int idx = threadIdx.x; idx < nl * Lf; idx += kFftThreads) {
      const float2 c = cmul(src[idx], p.bhat[idx & (Lf - 1)]);  // Lf: a power of two
      src[idx] = make_float2(c.x, -c.y);
    }
    run_stages(p, Lf, tw, nl, src, dst);
  }
  __syncthreads();

  // ---- store (crop, conjugation, scale and deconvolution) ----
  // with bin maxima: per (line, z bin) maxima in LDS (the free buffer; non-
  // negative floats order like their bit patterns), then one global atomic
  // maximum per (line, bin)
  uint32_t *bm = (uint32_t *)dst;
  const int nbz = p.bmax ? p.bm_nbin[2] : 0;
  if (p.bmax) {
    for (int i = threadIdx.x; i < nl * nbz; i += kFftThreads) bm[i] = 0u;
    __syncthreads();
  }
  uint32_t *bmx = p.bmax ? bm : nullptr;
  if (p.dec_out) store_lines<true>(p, lbase, nl, mag_nl, Lf, blue, src, bmx, nbz);
  else store_lines<false>(p, lbase, nl, mag_nl, Lf, blue, src, bmx, nbz);
  if (p.bmax) {
    __syncthreads();
    for (int i = threadIdx.x; i < nl * nbz; i += kFftThreads) {
      const int li = i / nbz, zb = i - li * nbz;
      const int64_t l = l0 + li, x = l / p.nb, y = l - x * p.nb;
      const int64_t bin = ((x / p.bm_bin[0]) * p.bm_nbin[1] + y / p.bm_bin[1]) * p.bm_nbin[2] + zb;
      atomicMax((uint32_t *)p.bmax + bin, bm[i]);
    }
  }
}

// Zero n complex values: 16-byte stores (two values) where the pointer is
// 16-byte aligned, the odd tail (and an unaligned start) element-wise.
__global__ void fill_kernel(float2 *p, int64_t n) {
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t head = ((uintptr_t)p & 15) ? 1 : 0;  // float2 is 8-byte aligned
  if (i0 == 0 && head && n > 0) p[0] = make_float2(0.f, 0.f);
  float4 *q = (float4 *)(p + head);
  const int64_t n2 = (n - head) / 2;
  for (int64_t i = i0; i < n2; i += st) q[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  if (i0 == 0 && n - head - 2 * n2 == 1) p[n - 1] = make_float2(0.f, 0.f);
}

} // namespace

// ceil(2^32 / d) for d > 1, 0 for d <= 1 (fdiv's multiplier)
static uint32_t div_magic(int64_t d) {
  return d > 1 ? (uint32_t)(((uint64_t)1 << 32) / (uint64_t)d + ((((uint64_t)1 << 32) % (uint64_t)d) ? 1 : 0)) : 0u;
}

hipError_t launch_fft_pass(const FftPass &pass, hipStream_t s) {
  if (pass.nlines <= 0) return hipSuccess;
  FftPass p = pass;
  p.mag_nin = div_magic(p.nin);
  p.mag_nout = div_magic(p.nout);
  p.mag_bmz = div_magic(p.bmax ? p.bm_bin[2] : 1);
  p.mag_lb = div_magic(p.lb);
  p.mag_last = div_magic(p.nlines - (p.nlines - 1) / p.lb * p.lb);
  p.use_mag_nb = p.nlines < ((int64_t)1 << 31) &&
                         (uint64_t)p.nlines * (uint64_t)p.nb < ((uint64_t)1 << 32) ? 1 : 0;
  p.mag_nb = div_magic(p.nb);
  {
    const int Lf0 = p.blue_n > 0 ? p.blue_n : p.L;
    int64_t ns = 1;
    for (int f = 0; f < p.nfac && f < 16; ++f) {
      p.mag_nbf[f] = div_magic(Lf0 / p.fac[f]);
      p.mag_ns[f] = div_magic(ns);
      p.twstep[f] = (int)(Lf0 / (ns * p.fac[f]));
      ns *= p.fac[f];
    }
  }
  int64_t blocks = (p.nlines + p.lb - 1) / p.lb;
  const int Lf = p.blue_n > 0 ? p.blue_n : p.L;
  size_t lds = (2 * (size_t)p.lb + 1) * Lf * sizeof(float2) + 2 * (size_t)p.lb * sizeof(int64_t);
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)fft_pass_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(fft_pass_kernel, dim3((unsigned)blocks), dim3(kFftThreads), lds, s, p);
  return hipGetLastError();
}

hipError_t launch_fill(float2 *p, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

} // namespace anufft
