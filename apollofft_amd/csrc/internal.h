// internal.h -- shared host/device definitions of libapollo_nufft_hip.
//
// Data layout in HBM (all row-major, z fastest, like ndarray Array3):
//   fine grid   float2[mx][my][mz]
//   modes       float2[nx][ny][nz]   (FFT order, k <= n/2 positive)
//   points      float coords[a][j*stride], float2 values[j]
// The kernel phi is evaluated from per-tap polynomials: for a point at grid
// coordinate t in cell c, the first tap is i0 = c - w + 1, x1 = i0 - t in
// [-w,-w+1], z = 2(x1 + w) - 1 in [-1,1], and tap i has phi(x1 + i) = P_i(z).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace anufft {

constexpr int kMinW = 2;
constexpr int kMaxW = 8;            // largest half-width with device kernels
constexpr int kMaxTaps = 2 * kMaxW; // non-zero taps per axis
constexpr int kPolyCoefs = 11;      // degree-10 per-tap polynomials
constexpr int kRadixBits = 8;       // LSD counting-sort digit width
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;       // points per thread per chunk
constexpr int kSortChunk = kSortThreads * kSortItems;
// One-pass binning (stable counting sort on the whole tile key) when the
// tile count fits the per-wave LDS count rows: kOneWaves waves per chunk,
// each ranking kOneSteps x 64 consecutive points.
constexpr int kOneWaves = 8;
constexpr int kOneThreads = 64 * kOneWaves;
constexpr int kOneSteps = 8;
constexpr int kOneChunk = kOneThreads * kOneSteps;
constexpr int kOneMaxBins = 9216;   // kOneWaves rows of 16-bit counts: 144 KiB of LDS
constexpr int kSpreadBinX = 4;      // spreading tile x-extent (a lane holds bin_x + 2w - 1 cells)
constexpr int kSpreadCols = 32;     // spreading tile (y,z) columns per axis, bin + 2w <= 32

// Spreading kernels: register columns on the vector ALUs, or FP16-split
// MFMA blocks (values of anufft_opts.kernel_method).
constexpr int kSpreadAuto = 0;
constexpr int kSpreadValu = 1;
constexpr int kSpreadMfma = 2;

// Per-tap kernel polynomials, passed by value (kernarg segment -> scalar regs).
// c[d][i] multiplies z^d for tap i; dc[d][i] is the same for d phi/dx.
struct KbPoly {
  float c[kPolyCoefs][kMaxTaps];
  float dc[kPolyCoefs][kMaxTaps];
};

// Geometry of the oversampled grid and the spreading tiles.
struct Geom {
  int m[3];          // oversampled sizes (1D: m[1] = m[2] = 1)
  int bin[3];        // tile size in fine cells
  int nbin[3];       // tiles per axis
  int pad[3];        // padded tile extents bin + 2w
  double scale[3];   // m / L: grid units per unit length
  double inv_m[3];   // 1 / m (rounded once)
  float inv_bin[3];  // 1 / bin (FP32): cell / bin = floor((cell + 0.5) inv_bin) for cells < 2^13
};

// cell / bin for 0 <= cell < 2^13, 1 <= bin < 2^13 without an integer
// division: (cell + 1/2) / bin lies at least 1/(2 bin) from the integers
// around it, far more than the FP32 rounding of the product (< 2^-10 / bin).
__host__ __device__ __forceinline__ int cell_div(int cell, const Geom &g, int a) {
  return (int)(((float)cell + 0.5f) * g.inv_bin[a]);
}

// Work of a subproblem for splitting the list among workgroups: one unit
// per kWorkUnit points plus kSubWork units of fixed cost (staging, scale,
// flush).
constexpr uint32_t kWorkUnit = 16;
constexpr uint32_t kSubWork = 48;

struct SortBuffers {
  uint32_t *keys[2];
  uint32_t *idx[2];
  uint32_t *cnt;       // per-chunk digit (one-pass: tile) counts, then offsets
  uint32_t *scan_tmp;  // digit (one-pass: tile) totals
  uint32_t *bin_start; // nbins + 1
  uint32_t *sub_ofs;   // nbins + 2; [nbins] = number of subproblems, [nbins + 1] = total work
  int4 *subs;          // (bin, start, count, work before it)
  int nbins;
  int nchunks;
  int passes;
  int max_sub;
  int sub_work;        // fixed cost of a subproblem in kWorkUnit units
};

// One batched FFT pass over lines of a 3D (or 1D) array.
struct FftPass {
  const float2 *in;
  float2 *out;
  int L;       // transform length
  int nin;     // inputs per line (nin < L: kept modes, zero padded)
  int nout;    // outputs per line (nout < L: kept modes only)
  int64_t nlines, nb;               // line l -> (a = l / nb, b = l % nb)
  int64_t in_sa, in_sb, in_si;      // element strides
  int64_t out_sa, out_sb, out_si;
  const double *dec_in;  // factor per input index (nin) or null
  const double *dec_out; // factor per output index (nout) or null
  double scale;
  int inverse;           // 0: exp(-i), 1: exp(+i) (unnormalized)
  const float2 *tw;      // exp(-2 pi i k / Lf), k in [0, Lf)
  int nfac;
  int fac[16];           // radices of Lf
  int lb;                // lines per block
  // Bluestein (lengths with a prime factor above kDirectPrimeMax): the
  // length-L DFT is a circular convolution of length blue_n (a power of two,
  // >= 2L - 1) done with two forward FFTs; Lf = blue_n, else Lf = L.
  int blue_n;
  const float2 *chirp;   // exp(-i pi n^2 / L), n in [0, L)
  const float2 *bhat;    // FFT_blue_n of the conjugate chirp filter, times 1/blue_n
  // Bin maxima of the written grid (the last pass of a 3D Type-2, lines
  // (a, b) = (x, y), output index = z): bmax[bin] = max |re|, |im| over the
  // bin's cells, with bins of bm_bin cells (bm_nbin per axis); zero[0 ..
  // zero_n) is cleared by the pass before (the first pass of the chain).
  float *bmax;
  int bm_bin[3];
  int bm_nbin[3];
  float *zero;
  int zero_n;
  // Division by the pass's runtime divisors as multiply-high by
  // ceil(2^32 / d) (0 for d = 1), set by launch_fft_pass: nin, nout, the
  // z bin width of the bin maxima, the lines of a full and of the last
  // block, and per stage f the butterflies per line Lf / fac[f] and the
  // stage span Ns (product of fac[0 .. f-1]); the twiddle step of stage f,
  // Lf / (Ns fac[f]).  mag_nb: the line index split l / nb, used when
  // nlines * nb < 2^32 (use_mag_nb), else the kernel divides in 64 bits.
  uint32_t mag_nin, mag_nout, mag_bmz, mag_lb, mag_last, mag_nb;
  int use_mag_nb;
  uint32_t mag_nbf[16], mag_ns[16];
  int twstep[16];
};

constexpr int kDirectPrimeMax = 16;   // larger prime radices go through Bluestein
constexpr int kBluesteinMaxN = 4096;  // largest Bluestein FFT (one line in LDS)

// Arguments shared by the 3D spread / interp kernels.
struct PointArgs {
  const float *x, *y, *z;
  int64_t stride;
  int64_t num_points;
  const uint32_t *perm;      // sorted order -> original point index
  const int4 *subs;
  const uint32_t *nsub;      // device: [0] subproblems, [1] their total work (subs[i].w units)
};

// Compact per-point records in sorted order for the MFMA kernels, 32 bytes
// per point as two float4: r[2s] = (z0, z1, z2, off), r[2s+1] = (c.x, c.y,
// key, idx).  z[a] = 1 - 2 (t - cell) of axis a (the argument of the tap
// polynomials, FP64-folded and rounded once, exactly as kb_taps computes
// it); off = x register cell | y column << 8 | z column << 16 of the first
// taps in the point's tile; c = value (Type-1, else 0); key = tile; idx =
// input index of the point.  off, key and idx are stored as their bits.
struct ZRec {
  float4 *r;
};

// Compact records built by the binning itself: the first digit pass reads
// the coordinates and values in input order and scatters each point's record
// (with its key and index) to its place; later passes move the records, and
// the last one writes them to `out` (tmp is the other buffer of the
// ping-pong).  The sorted keys and the permutation then live only in the
// records: PointArgs::perm is not written.
struct RecordOut {
  ZRec out;
  ZRec tmp;
  const float2 *values;  // point values (Type-1) or null
};

// One problem of the one-pass binning (sort.hip).  A single binning passes
// it by value; the batched binning reads a device table of them, chunk0 /
// scan0 being the problem's first chunk / scan block in the shared launches.
struct BinProblem {
  PointArgs pa;             // coordinates and point count
  Geom g;
  uint32_t *cnt;            // per-chunk tile counts, then offsets
  uint32_t *tot;            // tile totals
  uint32_t *bin_start;      // nbins + 1
  uint32_t *nsub;           // [0] subproblems, [1] their total work
  int4 *subs;
  ZRec out;                 // records (REC) ...
  uint32_t *keys_out, *idx_out;  // ... or sorted keys and the permutation
  const float2 *values;     // record values (Type-1 strengths, gradient u) or null
  int dim, nbins, bits, max_sub, sub_work, nchunks;
  uint32_t chunk0, scan0;
  int fused;                // one problem: the scatter launch also writes the list
};

// One problem of a batched spreading launch.
struct SpreadProblem {
  PointArgs pa;
  const float *rec;    // per-point records (launch_point_prep, vector-ALU kernel)
  const uint8_t *xo;   // per-point x offsets in the tile
  ZRec zr;             // compact records (built by sort_points, MFMA kernel)
  float2 *grid;
  Geom g;
};

constexpr int kBatchChunk = 16;  // descriptors per table-fill launch (kernarg size)
struct BatchChunk {
  int first;
  int count;
  SpreadProblem probs[kBatchChunk];
};

// ---- launchers (defined in the .hip files) ---------------------------------
hipError_t launch_fill(float2 *p, int64_t n, hipStream_t s);

hipError_t sort_points(const PointArgs &pa, const Geom &g, int dim, const SortBuffers &sb,
                       hipStream_t s, const RecordOut *rec = nullptr);
// Binning of nprob problems in shared launches (count, scan, subproblem
// lists, scatter: four launches for all), records written to each problem's
// out; every problem must take the one-pass binning.  host_probs' chunk0,
// scan0, bits and nchunks are filled here; dev_table holds nprob entries.
hipError_t sort_points_batched(BinProblem *host_probs, int nprob, BinProblem *dev_table,
                               hipStream_t s);
// Whether a tile count takes the one-pass binning, and the entries of
// SortBuffers::cnt the binning of num_points points needs.
bool sort_one_pass(int nbins);
size_t sort_count_entries(int64_t num_points, int nbins);

hipError_t launch_fft_pass(const FftPass &p, hipStream_t s);

// Per-point records in sorted order for the vector-ALU spreading kernel:
// point_rec_floats(w) floats (value and the tap weights of the three axes
// with their tile offsets) and one x-offset byte per point, filled by
// launch_point_prep (values may be NULL for Type-2).
int point_rec_floats(int w);
hipError_t launch_point_prep(int w, const PointArgs &pa, const float2 *values, const Geom &g,
                             const KbPoly &poly, float *rec, uint8_t *xo, hipStream_t s);
hipError_t launch_spread3d(int w, int method, const PointArgs &pa, const float *rec,
                           const uint8_t *xo, ZRec zr, float2 *grid, const Geom &g,
                           const KbPoly &poly, const KbPoly *dpoly, int num_blocks, hipStream_t s);
hipError_t launch_spread3d_batched(int w, int method, const SpreadProblem *host_probs, int nprob,
                                   SpreadProblem *dev_table, uint32_t *dev_prefix,
                                   const KbPoly &poly, const KbPoly *dpoly, int num_blocks,
                                   hipStream_t s);
// MFMA interpolation needs 2 nbins floats of scratch (bin and tile maxima).
// Tile maxima of the grid for the MFMA interpolation (no-op for the vector
// ALU method); launch_interp3d reads them from the same scratch.
// bin_max_done: the bin maxima (first half of the scratch) were written by the
// last FFT pass (FftPass::bmax), so only the tile maxima are computed.
hipError_t launch_tile_max3d(int w, int method, const float2 *grid, float *scratch, const Geom &g,
                             hipStream_t s, bool bin_max_done = false);
// cus: the device's CU count (the windowed kernel weights the shares of the
// older and younger workgroup of a CU when the grid holds two per CU).
hipError_t launch_interp3d(int w, int method, const PointArgs &pa, ZRec zr, const float2 *grid,
                           float *scratch, float2 *out, const Geom &g, const KbPoly &poly,
                           const KbPoly *dpoly, int num_blocks, int cus, hipStream_t s);
hipError_t launch_grad3d(int w, int method, const PointArgs &pa, ZRec zr, const float2 *grid,
                         const float2 *u, float *gx, float *gy, float *gz, const Geom &g,
                         const KbPoly &poly, const KbPoly *dpoly, int num_blocks, hipStream_t s);

// Direct FP64 sums (direct.hip): type 1 writes every mode, type 2 every point.
hipError_t launch_direct(int type, const PointArgs &pa, int dim, const int64_t n[3],
                         const double L[3], const float2 *in, double2 *out, hipStream_t s);
// Type-1 at nsel selected flat mode indices, complex128 values, FP64.
hipError_t launch_direct_select(const PointArgs &pa, int dim, const int64_t n[3],
                                const double L[3], const double2 *values, const int64_t *sel,
                                int64_t nsel, double2 *out, hipStream_t s);
// The same sums over FP64 coordinates (stride as PointArgs) and complex128 data.
hipError_t launch_direct_f64(int type, const double *x, const double *y, const double *z,
                             int64_t stride, int64_t num_points, int dim, const int64_t n[3],
                             const double L[3], const double2 *in, double2 *out, hipStream_t s);

// 1D on binned tiles (records from sort_points with dim 1, values = the
// point values for spreading, u for the gradient): spreading adds to grid;
// interpolation writes out (gx NULL) or the coordinate gradient gx.
hipError_t launch_spread1d(int w, const PointArgs &pa, ZRec zr, float2 *grid, const Geom &g,
                           const KbPoly &poly, int num_blocks, hipStream_t s);
hipError_t launch_interp1d(int w, const PointArgs &pa, ZRec zr, const float2 *grid, float2 *out,
                           float *gx, const Geom &g, const KbPoly &poly, int num_blocks,
                           hipStream_t s);
// Interior (cells) of the 1D tiles: a padded tile is one wave of 64 cells.
constexpr int kTile1d = 64;
inline int bin1d(int w) { return kTile1d - 2 * w + 1; }

// LDS bytes of the 3D spread/interp kernels for a geometry.
size_t spread_lds_bytes(int w, int method, const Geom &g);
int spread_bin_x(int w, int method);        // spreading tile interior along x
int interp_blocks_per_cu(int w, int method);  // workgroups per CU of 3D interpolation
// The MFMA gradient kernel keeps two x cells per wave; at w = 8 (32-cell
// tiles, 16 waves) it would not fit the registers, so w = 8 uses LDS tiles.
inline bool grad_uses_mfma(int w, int method) { return method == kSpreadMfma && w < 8; }
int spread_block_threads(int w, int method);
size_t interp_lds_bytes(int w, const Geom &g, bool grad);

// ---- device helpers ----------------------------------------------------------
#if defined(__HIPCC__)
// Fold a coordinate into [0, m) grid units and return the cell index.
// t = x * scale, q = floor(t * (1/m)), t -= q m, then t += m if negative or
// t -= m if t >= m (q may be off by one near a multiple of m: no FP64
// division); cell = min(floor(t), m-1).  q*m is exact (integer times integer
// below 2^53), so the subtraction rounds once whether or not it is
// contracted into an FMA; the CPU replay of the binning repeats this
// sequence in IEEE double and gets identical cells.
__device__ __forceinline__ double fold_coord(float xf, double scale, int m, double inv_m,
                                             int &cell) {
  double t = (double)xf * scale;
  double md = (double)m;
  double q = floor(t * inv_m);
  t = t - q * md;
  if (t < 0.0) t = t + md;
  else if (t >= md) t = t - md;
  int c = (int)floor(t);
  if (c > m - 1) c = m - 1;
  if (c < 0) c = 0;
  cell = c;
  return t;
}

// Compact record fields of point j (the ZRec layout): polynomial arguments
// z[a] = 1 - 2 (t - cell) and the tile offsets of the first taps (x: cell mod
// bin; y, z: cell mod bin + 1, the tile starting w cells before its bin).
__device__ __forceinline__ void point_zrec(const PointArgs &pa, const Geom &g, int64_t j,
                                           float (&z)[3], int &off) {
  off = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float *cp = a == 0 ? pa.x : (a == 1 ? pa.y : pa.z);
    int cell;
    const double t = fold_coord(cp[j * pa.stride], g.scale[a], g.m[a], g.inv_m[a], cell);
    z[a] = (float)(1.0 - 2.0 * (t - (double)cell));
    const int o = cell - (cell / g.bin[a]) * g.bin[a];
    off |= (a == 0 ? o : o + 1) << (8 * a);
  }
}

// Evaluate the 2w tap weights of one axis for grid coordinate t in cell c
// (t in [c, c+1]).  The taps are i0 + i, i0 = c - w + 1, i in [0, 2w), at
// distances x1 + i with x1 = i0 - t in [-w, -w+1]; with z = 2(x1 + w) - 1 =
// 1 - 2(t - c) the weight of tap i is P_i(z).  These are exactly the
// non-zero taps of the reference loop p in [-w, w] around round(t)
// (dimension_3d.rs:294-305): a tap at distance +-w has phi = 0
// (kaiser_bessel.rs:111-112), which the two endpoint cases enforce.  All
// points of one cell share i0.  Returns i0.
template <int W>
__device__ __forceinline__ int kb_taps(double t, int cell, const KbPoly &P, float *wt) {
  const float z = (float)(1.0 - 2.0 * (t - (double)cell));
#pragma unroll
  for (int i = 0; i < 2 * W; ++i) {
    float v = P.c[kPolyCoefs - 1][i];
#pragma unroll
    for (int d = kPolyCoefs - 2; d >= 0; --d) v = fmaf(v, z, P.c[d][i]);
    wt[i] = v;
  }
  if (z <= -1.0f) wt[0] = 0.0f;          // t = c + 1: first tap at distance -w
  if (z >= 1.0f) wt[2 * W - 1] = 0.0f;   // t = c: last tap at distance +w
  return cell - W + 1;
}

// Value and derivative (d phi / d x, per grid unit) taps.
template <int W>
__device__ __forceinline__ int kb_taps_deriv(double t, int cell, const KbPoly &P, float *wt,
                                             float *dwt) {
  const float z = (float)(1.0 - 2.0 * (t - (double)cell));
#pragma unroll
  for (int i = 0; i < 2 * W; ++i) {
    float v = P.c[kPolyCoefs - 1][i];
    float dv = P.dc[kPolyCoefs - 1][i];
#pragma unroll
    for (int d = kPolyCoefs - 2; d >= 0; --d) {
      v = fmaf(v, z, P.c[d][i]);
      dv = fmaf(dv, z, P.dc[d][i]);
    }
    wt[i] = v;
    dwt[i] = dv;
  }
  if (z <= -1.0f) {
    wt[0] = 0.0f;
    dwt[0] = 0.0f;
  }
  if (z >= 1.0f) {
    wt[2 * W - 1] = 0.0f;
    dwt[2 * W - 1] = 0.0f;
  }
  return cell - W + 1;
}
#endif

} // namespace anufft
