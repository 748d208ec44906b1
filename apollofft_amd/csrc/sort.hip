// sort.hip -- point binning by a stable LSD counting sort on integer tile keys.
//
// Replaces the comparison sorts sort_points_{1d,3d} / sort_positions_{1d,3d}
// of apollo-nufft (transform/dimension_1d.rs:54-99, dimension_3d.rs:74-161):
// the reference orders points by bucket with float tie-breaks; here the key
// is the spreading tile of the point and ties keep the input order, so the
// permutation is a deterministic function of the coordinates that a CPU
// replay reproduces bit for bit.
//
// Pipeline (one stream, no host synchronisation):
//   keys   : key[j] = tile of point j, per-chunk histogram of digit 0
//   per digit pass: per-digit scan of the chunk counts -> stable scatter
//                   (histogram of the next digit)
//   bounds : bin_start[b] = lower_bound(sorted keys, b)
//   subproblems (one block): scan of per-tile subproblem counts, list
//   subs   : subproblems of at most max_sub points per tile, tiles in pencil
//            order (x fastest)
#include <algorithm>

#include "internal.h"

namespace anufft {
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t tile_key(const PointArgs &pa, const Geom &g, int dim,
                                             int64_t j) {
  int cx, cy = 0, cz = 0;
  fold_coord(pa.x[j * pa.stride], g.scale[0], g.m[0], g.inv_m[0], cx);
  if (dim == 3) {
    fold_coord(pa.y[j * pa.stride], g.scale[1], g.m[1], g.inv_m[1], cy);
    fold_coord(pa.z[j * pa.stride], g.scale[2], g.m[2], g.inv_m[2], cz);
  }
  uint32_t bx = (uint32_t)(cx / g.bin[0]);
  uint32_t by = (uint32_t)(cy / g.bin[1]);
  uint32_t bz = (uint32_t)(cz / g.bin[2]);
  return (bx * (uint32_t)g.nbin[1] + by) * (uint32_t)g.nbin[2] + bz;
}

// keys + digit-0 histogram of each chunk, written digit-major.
__global__ __launch_bounds__(kSortThreads) void keys_kernel(PointArgs pa, Geom g, int dim,
                                                            uint32_t *keys, uint32_t *cnt,
                                                            int nchunks) {
  __shared__ uint32_t hist[kRadix];
  for (int d = threadIdx.x; d < kRadix; d += kSortThreads) hist[d] = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * kSortChunk;
  for (int i = threadIdx.x; i < kSortChunk; i += kSortThreads) {
    int64_t j = base + i;
    if (j < pa.num_points) {
      uint32_t k = tile_key(pa, g, dim, j);
      keys[j] = k;
      atomicAdd(&hist[k & (kRadix - 1)], 1u);
    }
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kRadix; d += kSortThreads)
    cnt[(int64_t)d * nchunks + blockIdx.x] = hist[d];
}

// Keys of the current order: a key array, or the key field of the records.
struct KeySrc {
  const uint32_t *k;
  const float4 *r;
  __device__ __forceinline__ uint32_t operator[](int64_t j) const {
    return r ? __float_as_uint(r[2 * j + 1].z) : k[j];
  }
};

// Per-chunk histogram of digit `shift` (passes after the first).
__global__ __launch_bounds__(kSortThreads) void hist_kernel(KeySrc keys, int64_t n, int shift,
                                                            uint32_t *cnt, int nchunks) {
  __shared__ uint32_t hist[kRadix];
  for (int d = threadIdx.x; d < kRadix; d += kSortThreads) hist[d] = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * kSortChunk;
  for (int i = threadIdx.x; i < kSortChunk; i += kSortThreads) {
    int64_t j = base + i;
    if (j < n) atomicAdd(&hist[(keys[j] >> shift) & (kRadix - 1)], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kRadix; d += kSortThreads)
    cnt[(int64_t)d * nchunks + blockIdx.x] = hist[d];
}

// ---- exclusive scan of a uint32 array (reduce, scan block sums, apply) ----
// Exclusive scan over a block of NT threads (NT / 64 waves): wave scans by
// shuffles, wave totals through LDS (sh: NT / 64 entries); two barriers.
template <int NT>
__device__ __forceinline__ uint32_t block_scan_waves(uint32_t v, uint32_t *sh, uint32_t &total) {
  constexpr int NWV = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[wave] = x;
  __syncthreads();
  uint32_t before = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < NWV; ++w) {
    const uint32_t t = sh[w];
    before += w < wave ? t : 0u;
    total += t;
  }
  __syncthreads();  // sh may be reused
  return before + x - v;
}

// Exclusive scan of each digit's row of chunk counts, in place: one block per
// digit, the row walked in tiles of kScanTile with a running carry;
// tot[d] = the row's total.  The scatter adds the digit's base (the totals
// of the smaller digits) itself, so one launch replaces a whole-array scan.
__global__ __launch_bounds__(kScanThreads) void scan_rows_kernel(uint32_t *cnt, int nchunks,
                                                                 uint32_t *tot) {
  __shared__ uint32_t sh[kScanThreads / 64];
  uint32_t *row = cnt + (int64_t)blockIdx.x * nchunks;
  uint32_t carry = 0;
  for (int base = 0; base < nchunks; base += kScanTile) {
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      const int j = base + threadIdx.x * kScanItems + i;
      v[i] = j < nchunks ? row[j] : 0u;
      s += v[i];
    }
    uint32_t total;
    uint32_t ex = block_scan_waves<kScanThreads>(s, sh, total) + carry;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      const int j = base + threadIdx.x * kScanItems + i;
      if (j < nchunks) row[j] = ex;
      ex += v[i];
    }
    carry += total;
  }
  if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

// ---- stable scatter of one digit --------------------------------------------
// One block per chunk; the chunk is walked in input order, kSortThreads points
// per step.  Ranks inside a wave come from matching digits with per-bit
// ballots; ranks across the waves of a step from per-wave digit counts.
// With REC the pass moves 32-byte records instead of separate keys and
// indices: the first pass builds each point's record from its coordinates
// and value (keys_in holds the keys of the input order), later passes read
// the records of the current order from `in`; all write to `out`.
struct RecMove {
  ZRec in, out;
  const float2 *values;  // value field present (Type-1) or null
  PointArgs pa;
  Geom g;
  int first;
};

template <bool REC>
__global__ __launch_bounds__(kSortThreads) void scatter_kernel(
    const uint32_t *keys_in, const uint32_t *idx_in, uint32_t *keys_out, uint32_t *idx_out,
    int64_t n, int shift, const uint32_t *offs, const uint32_t *tot, int nchunks, RecMove rm) {
  constexpr int kWaves = kSortThreads / 64;
  static_assert(kSortThreads == kRadix, "one thread per digit");
  __shared__ uint32_t running[kRadix];
  __shared__ uint32_t wcount[kWaves][kRadix];
  __shared__ uint32_t sh[kWaves];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  {
    // start of digit tid in this chunk: digits below it, then its earlier chunks
    uint32_t all;
    const uint32_t dbase = block_scan_waves<kSortThreads>(tot[tid], sh, all);
    running[tid] = dbase + offs[(int64_t)tid * nchunks + blockIdx.x];
    for (int w = 0; w < kWaves; ++w) wcount[w][tid] = 0;
  }
  __syncthreads();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int64_t base = (int64_t)blockIdx.x * kSortChunk;
  for (int step = 0; step < kSortItems; ++step) {
    int64_t j = base + (int64_t)step * kSortThreads + tid;
    bool valid = j < n;
    uint32_t key = 0u, idx = 0u;
    float4 ra = make_float4(0.f, 0.f, 0.f, 0.f), rb = ra;
    if (!REC) {
      key = valid ? keys_in[j] : 0u;
      idx = valid ? (idx_in ? idx_in[j] : (uint32_t)j) : 0u;
    } else if (valid) {
      if (rm.first) {
        key = keys_in[j];
        idx = (uint32_t)j;
        float z[3];
        int off;
        point_zrec(rm.pa, rm.g, j, z, off);
        const float2 c = rm.values ? rm.values[j] : make_float2(0.f, 0.f);
        ra = make_float4(z[0], z[1], z[2], __int_as_float(off));
        rb = make_float4(c.x, c.y, __uint_as_float(key), __uint_as_float(idx));
      } else {
        ra = rm.in.r[2 * j];
        rb = rm.in.r[2 * j + 1];
        key = __float_as_uint(rb.z);
      }
    }
    uint32_t d = (key >> shift) & (kRadix - 1);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kRadixBits; ++b) {
      bool bit = (d >> b) & 1u;
      uint64_t m = __ballot(valid && bit);
      peers &= bit ? m : ~m;
    }
    uint32_t rank = (uint32_t)__popcll(peers & lt);
    if (valid && (peers & lt) == 0ull) wcount[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = running[d] + rank;
      for (int w = 0; w < wave; ++w) pos += wcount[w][d];
      if (REC) {
        rm.out.r[2 * (int64_t)pos] = ra;
        rm.out.r[2 * (int64_t)pos + 1] = rb;
      } else {
        keys_out[pos] = key;
        idx_out[pos] = idx;
      }
    }
    __syncthreads();
    for (int dd = tid; dd < kRadix; dd += kSortThreads) {
      uint32_t s = 0;
      for (int w = 0; w < kWaves; ++w) {
        s += wcount[w][dd];
        wcount[w][dd] = 0;
      }
      running[dd] += s;
    }
    __syncthreads();
  }
}

// bin_start[b] = first sorted position with key >= b, b in [0, nbins].
__global__ void bounds_kernel(KeySrc keys, int64_t n, uint32_t *bin_start, int nbins) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nbins) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < (uint32_t)b) lo = mid + 1;
    else hi = mid;
  }
  bin_start[b] = (uint32_t)lo;
}

// The subproblem list runs over the tiles in pencil order -- x fastest,
// then z, then y -- so that consecutive entries are neighbouring tiles
// along x; q is the position of a tile in that order.
__device__ __forceinline__ int pencil_bin(int q, int3 nb) {
  const int bx = q % nb.x, pz = q / nb.x;
  return (bx * nb.y + pz / nb.z) * nb.z + pz % nb.z;
}

// 64-bit exclusive scan over a block of NT threads (see block_scan_waves).
template <int NT>
__device__ __forceinline__ uint64_t block_scan_waves64(uint64_t v, uint64_t *sh, uint64_t &total) {
  constexpr int NWV = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[wave] = x;
  __syncthreads();
  uint64_t before = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < NWV; ++w) {
    const uint64_t t = sh[w];
    before += w < wave ? t : 0ull;
    total += t;
  }
  __syncthreads();  // sh may be reused
  return before + x - v;
}

// Subproblem list in one block: tiles in pencil order, ceil(count / max_sub)
// subproblems each, exclusive scan of those counts in rounds of kSubRound
// tiles (kSubPer consecutive tiles per thread).  subs[i].w = work before
// subproblem i in list order, in units of kWorkUnit points, a subproblem
// weighing exactly ceil(count / kWorkUnit) + sub_work (its fixed cost); nsub[0] =
// number of subproblems, nsub[1] = total work.  The subproblem counts and
// the work are scanned together (one 64-bit scan: counts in the high word).
// The entries of a round are written by the whole block, each finding its
// tile by a binary search of the tiles' first entries (a tile of a dense
// cluster can hold hundreds of subproblems).
// With from_tot (the tile totals tot of the one-pass binning) the kernel
// first writes bin_start[b] = sum of tot[d] over d < b, b in [0, nbins];
// tot == NULL with no points means all tiles are empty.
constexpr int kSubThreads = 1024;
constexpr int kSubPer = 4;
// LDS of the list for a block of NT threads (rounds of NT * kSubPer tiles).
template <int NT>
struct SubLds {
  static constexpr int kRound = NT * kSubPer;
  uint32_t sh[NT / 64];
  uint64_t sh64[NT / 64];
  uint32_t first[kRound];  // first entry of each tile of the round
  int4 tinfo[kRound];      // tile: bin, first point, count, work before it
};
template <int NT>
__device__ __forceinline__ void subproblems_body(SubLds<NT> &L, uint32_t *bin_start,
                                                 const uint32_t *tot, int from_tot, int nbins,
                                                 int3 nb, int max_sub, int sub_work,
                                                 uint32_t *nsub, int4 *subs) {
  constexpr int kSubRound = SubLds<NT>::kRound;
  uint32_t *sh = L.sh;
  uint64_t *sh64 = L.sh64;
  uint32_t *first = L.first;
  int4 *tinfo = L.tinfo;
  if (from_tot) {
    uint32_t run = 0;
    for (int base = 0; base <= nbins; base += kSubRound) {
      const int b0 = base + threadIdx.x * kSubPer;
      uint32_t v[kSubPer], sum = 0;
#pragma unroll
      for (int i = 0; i < kSubPer; ++i) {
        v[i] = (tot && b0 + i < nbins) ? tot[b0 + i] : 0u;
        sum += v[i];
      }
      uint32_t total;
      uint32_t ex = run + block_scan_waves<NT>(sum, sh, total);
#pragma unroll
      for (int i = 0; i < kSubPer; ++i) {
        if (b0 + i <= nbins) bin_start[b0 + i] = ex;
        ex += v[i];
      }
      run += total;
    }
    __syncthreads();  // bin_start is read below by other threads of the block
  }
  const uint32_t ms = (uint32_t)max_sub;
  const uint32_t sw = (uint32_t)sub_work;
  const uint32_t wfull = (ms + kWorkUnit - 1) / kWorkUnit + sw;  // a full subproblem
  uint32_t carry = 0, carry_w = 0;
  for (int base = 0; base < nbins; base += kSubRound) {
    const int q0 = base + threadIdx.x * kSubPer;
    int b[kSubPer];
    uint32_t start[kSubPer], c[kSubPer], k[kSubPer], wk[kSubPer];
    uint64_t sum = 0;
#pragma unroll
    for (int i = 0; i < kSubPer; ++i) {
      b[i] = 0;
      start[i] = c[i] = k[i] = wk[i] = 0u;
      if (q0 + i < nbins) {
        b[i] = pencil_bin(q0 + i, nb);
        start[i] = bin_start[b[i]];
        c[i] = bin_start[b[i] + 1] - start[i];
        k[i] = (c[i] + ms - 1) / ms;
        // k - 1 full subproblems and the rest: each interval of the work
        // prefix is its fixed cost plus its own units, for any max_sub
        const uint32_t rest = k[i] > 0 ? c[i] - (k[i] - 1) * ms : 0u;
        wk[i] = k[i] > 0 ? (k[i] - 1) * wfull + (rest + kWorkUnit - 1) / kWorkUnit + sw : 0u;
      }
      sum += ((uint64_t)k[i] << 32) | wk[i];
    }
    uint64_t total64;
    const uint64_t ex = block_scan_waves64<NT>(sum, sh64, total64);
    uint32_t o = (uint32_t)(ex >> 32), w = carry_w + (uint32_t)ex;
#pragma unroll
    for (int i = 0; i < kSubPer; ++i) {
      first[threadIdx.x * kSubPer + i] = o;
      tinfo[threadIdx.x * kSubPer + i] = make_int4(b[i], (int)start[i], (int)c[i], (int)w);
      o += k[i];
      w += wk[i];
    }
    __syncthreads();
    const uint32_t total = (uint32_t)(total64 >> 32), total_w = (uint32_t)total64;
    for (uint32_t t = threadIdx.x; t < total; t += NT) {
      // owner = last tile with first <= t; it has t - first < its count
      int lo = 0, hi = kSubRound - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= t) lo = mid;
        else hi = mid - 1;
      }
      const int4 ti = tinfo[lo];
      const uint32_t tl = t - first[lo];
      const uint32_t cnt = min((uint32_t)ti.z - tl * ms, ms);
      subs[carry + t] = make_int4(ti.x, (int)((uint32_t)ti.y + tl * ms), (int)cnt,
                                  (int)((uint32_t)ti.w + tl * wfull));
    }
    carry += total;
    carry_w += total_w;
  }
  if (threadIdx.x == 0) {
    nsub[0] = carry;
    nsub[1] = carry_w;
  }
}

__global__ __launch_bounds__(kSubThreads) void subproblems_kernel(uint32_t *bin_start,
                                                                  const uint32_t *tot, int from_tot,
                                                                  int nbins, int3 nb, int max_sub,
                                                                  int sub_work, uint32_t *nsub,
                                                                  int4 *subs) {
  __shared__ SubLds<kSubThreads> L;
  subproblems_body<kSubThreads>(L, bin_start, tot, from_tot, nbins, nb, max_sub, sub_work, nsub,
                                subs);
}

// ---- one-pass binning (tile count <= kOneMaxBins) ---------------------------
// A stable counting sort on the whole tile key in one scatter: chunks of
// kOneChunk points in input order, each split into kOneWaves contiguous runs
// of kOneSteps x 64 points, one run per wave.
//   chunk_count   : per-chunk tile histogram, cnt[chunk][tile]
//   chunk_scan    : per-tile exclusive scan over the chunks (in place) and
//                   the tile totals
//   subproblems   : bin_start = exclusive scan of the totals, and the list
//                   (up to kFusedMaxBins tiles: block 0 of the scatter
//                   launch, the chunks then scan the totals themselves)
//   chunk_scatter : each wave ranks its run's points among equal tiles with
//                   ballots (earlier lanes, then earlier steps through a
//                   per-wave LDS row of counts); the rows are scanned across
//                   the waves; a point's place is bin_start[tile] + its
//                   chunk's offset + its wave's offset + its rank.
// Ties keep input order, so the result equals the LSD passes' (and
// numpy.argsort(keys, kind="stable")).

// Peers of this lane (lanes of the wave with the same key among `valid`
// lanes), from one ballot per key bit.
__device__ __forceinline__ uint64_t key_peers(uint32_t key, bool valid, int bits) {
  uint64_t peers = __ballot(valid);
  for (int b = 0; b < bits; ++b) {
    const bool bit = (key >> b) & 1u;
    const uint64_t m = __ballot(valid && bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

// Coordinates of point j (zeros past the end; dim 1 keeps x only).  The
// loads are unconditional, from point 0 past the end (callers launch only
// with points; 1D point arguments alias y and z to x), so that no branch
// around a load makes the compiler wait for it there: all loads of a
// thread stay in flight together.
__device__ __forceinline__ void load_point(const PointArgs &pa, int dim, int64_t j, bool valid,
                                           float (&c)[3]) {
  const int64_t o = (valid ? j : 0) * pa.stride;
  const float x = pa.x[o], y = pa.y[o], z = pa.z[o];
  c[0] = valid ? x : 0.f;
  c[1] = (valid && dim == 3) ? y : 0.f;
  c[2] = (valid && dim == 3) ? z : 0.f;
}

// Tile key of a point from its coordinates (tile_key's arithmetic); with
// REC also the record fields of point_zrec (polynomial arguments, offsets).
template <bool REC>
__device__ __forceinline__ uint32_t fold_point(const float (&c)[3], const Geom &g, int dim,
                                               float (&z)[3], int &off) {
  int cell[3] = {0, 0, 0};
  off = 0;
  z[0] = z[1] = z[2] = 0.f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (a > 0 && dim != 3) break;
    const double t = fold_coord(c[a], g.scale[a], g.m[a], g.inv_m[a], cell[a]);
    if (REC) z[a] = (float)(1.0 - 2.0 * (t - (double)cell[a]));
  }
  int b[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    b[a] = (a > 0 && dim != 3) ? 0 : cell_div(cell[a], g, a);
    if (REC && (a == 0 || dim == 3)) {
      const int o = cell[a] - b[a] * g.bin[a];
      off |= (a == 0 ? o : o + 1) << (8 * a);
    }
  }
  return ((uint32_t)b[0] * (uint32_t)g.nbin[1] + (uint32_t)b[1]) * (uint32_t)g.nbin[2] + (uint32_t)b[2];
}

// Tile histogram of one chunk: every thread loads its points' coordinates
// first (all loads in flight at once), then folds them and counts each in
// LDS (the sums do not depend on the order of the atomic adds).
constexpr int kCountThreads = 1024;
constexpr int kCountPer = kOneChunk / kCountThreads;
// The problem of this block: the one passed by value, or (BATCH) the last
// table entry whose first block (chunk0 or scan0) is at most blockIdx.x.
template <bool BATCH, bool SCAN>
__device__ __forceinline__ const BinProblem &bin_problem(const BinProblem &one,
                                                          const BinProblem *table, int nprob,
                                                          uint32_t &local) {
  if (!BATCH) {
    local = blockIdx.x;
    return one;
  }
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((SCAN ? table[mid].scan0 : table[mid].chunk0) <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  local = blockIdx.x - (SCAN ? table[lo].scan0 : table[lo].chunk0);
  return table[lo];
}

template <bool BATCH>
__global__ __launch_bounds__(kCountThreads) void chunk_count_kernel(BinProblem one,
                                                                    const BinProblem *table,
                                                                    int nprob) {
  extern __shared__ uint32_t hist[];  // nbins
  uint32_t chunk;
  const BinProblem &pr = bin_problem<BATCH, false>(one, table, nprob, chunk);
  const int nbins = pr.nbins, dim = pr.dim;
  const PointArgs &pa = pr.pa;
  for (int d = threadIdx.x; d < nbins; d += kCountThreads) hist[d] = 0;
  const int64_t base = (int64_t)chunk * kOneChunk + threadIdx.x;
  float c[kCountPer][3];
#pragma unroll
  for (int s = 0; s < kCountPer; ++s) {
    const int64_t j = base + (int64_t)s * kCountThreads;
    load_point(pa, dim, j, j < pa.num_points, c[s]);
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < kCountPer; ++s) {
    const int64_t j = base + (int64_t)s * kCountThreads;
    const bool valid = j < pa.num_points;
    float z[3];
    int off;
    if (valid) atomicAdd(&hist[fold_point<false>(c[s], pr.g, dim, z, off)], 1u);
  }
  __syncthreads();
  uint32_t *row = pr.cnt + (int64_t)chunk * nbins;
  for (int d = threadIdx.x; d < nbins; d += kCountThreads) row[d] = hist[d];
}

// cnt[c][d] -> sum of cnt[c'][d] over c' < c (in place); tot[d] = the total.
// A block takes 16 tiles; its 1024 threads split the chunks into 64
// segments per tile (lane group of 16 = 16 consecutive tiles of one chunk),
// so each thread has one short run of loads in flight.
constexpr int kScanDigits = 16;
constexpr int kChunkScanSegs = 64;
template <bool BATCH>
__global__ __launch_bounds__(kScanDigits * kChunkScanSegs) void chunk_scan_kernel(
    BinProblem one, const BinProblem *table, int nprob) {
  __shared__ uint32_t seg[kChunkScanSegs][kScanDigits];
  uint32_t blk;
  const BinProblem &pr = bin_problem<BATCH, true>(one, table, nprob, blk);
  uint32_t *cnt = pr.cnt;
  const int nchunks = pr.nchunks, nbins = pr.nbins;
  const int dl = threadIdx.x & (kScanDigits - 1), sg = threadIdx.x / kScanDigits;
  const int d = (int)blk * kScanDigits + dl;
  const int per = (nchunks + kChunkScanSegs - 1) / kChunkScanSegs;
  const int c0 = min(nchunks, sg * per), c1 = min(nchunks, c0 + per);
  uint32_t s = 0;
  if (d < nbins) {
#pragma unroll 8
    for (int c = c0; c < c1; ++c) s += cnt[(int64_t)c * nbins + d];
  }
  seg[sg][dl] = s;
  __syncthreads();
  uint32_t before = 0, total = 0;
#pragma unroll 16
  for (int k = 0; k < kChunkScanSegs; ++k) {
    const uint32_t v = seg[k][dl];
    before += k < sg ? v : 0u;
    total += v;
  }
  if (d < nbins) {
#pragma unroll 8
    for (int c = c0; c < c1; ++c) {
      uint32_t *q = cnt + (int64_t)c * nbins + d;
      const uint32_t v = *q;
      *q = before;
      before += v;
    }
    if (sg == 0) pr.tot[d] = total;
  }
}

// Subproblem lists of a batch: one block per problem.
__global__ __launch_bounds__(kSubThreads) void subproblems_batched_kernel(const BinProblem *table) {
  const BinProblem &pr = table[blockIdx.x];
  const int3 nb = pr.dim == 3 ? make_int3(pr.g.nbin[0], pr.g.nbin[1], pr.g.nbin[2])
                              : make_int3(1, 1, 1);
  __shared__ SubLds<kSubThreads> L;
  subproblems_body<kSubThreads>(L, pr.bin_start, pr.pa.num_points > 0 ? pr.tot : nullptr, 1,
                                pr.nbins, nb, pr.max_sub, pr.sub_work, pr.nsub, pr.subs);
}

// Fused list (one problem, one.fused): block 0 writes bin_start and the
// subproblem list from the tile totals, beside the scatter; every other
// block scans the totals into LDS for the tile starts of its own points.
// Otherwise bin_start comes from the list launch before this one.
// Dynamic LDS: kOneWaves rows of nbins 16-bit counts, then (fused) nbins
// 32-bit tile starts; block 0 uses it as a SubLds<kOneThreads>.
constexpr int kFusedTotPer = 16;  // tile totals per thread in the fused scan
// Largest tile count of the fused list: 16 + 4 bytes of LDS per tile within
// the 160 KiB of a workgroup (and kFusedTotPer totals per thread).
constexpr int kFusedMaxBins = 8176;
static_assert(kFusedMaxBins <= kFusedTotPer * kOneThreads, "fused scan covers the tiles");
static_assert(20 * kFusedMaxBins + 4 * kOneWaves <= 160 * 1024, "fused scatter LDS");
static_assert(sizeof(SubLds<kOneThreads>) + 4 * kOneWaves <= 160 * 1024, "list LDS");
template <bool REC, bool BATCH>
__global__ __launch_bounds__(kOneThreads) void chunk_scatter_kernel(BinProblem one,
                                                                    const BinProblem *table,
                                                                    int nprob) {
  extern __shared__ uint4 scatter_lds[];
  __shared__ uint32_t wsum[kOneWaves];
  uint16_t *rows = (uint16_t *)scatter_lds;
  const bool fused = !BATCH && one.fused;
  if (fused && blockIdx.x == 0) {
    const int3 nb = one.dim == 3 ? make_int3(one.g.nbin[0], one.g.nbin[1], one.g.nbin[2])
                                 : make_int3(1, 1, 1);
    subproblems_body<kOneThreads>(*(SubLds<kOneThreads> *)scatter_lds, one.bin_start, one.tot, 1,
                                  one.nbins, nb, one.max_sub, one.sub_work, one.nsub, one.subs);
    return;
  }
  uint32_t chunk;
  const BinProblem &pr = bin_problem<BATCH, false>(one, table, nprob, chunk);
  if (fused) chunk -= 1;
  const int nbins = pr.nbins, dim = pr.dim, bits = pr.bits;
  const PointArgs &pa = pr.pa;
  for (int i = threadIdx.x; i < kOneWaves / 2 * nbins; i += kOneThreads) ((uint32_t *)rows)[i] = 0u;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint16_t *row = rows + wave * nbins;
  uint32_t *bst = (uint32_t *)(rows + kOneWaves * nbins);
  const int64_t j0 = (int64_t)chunk * kOneChunk + (int64_t)wave * (64 * kOneSteps) + lane;
  // the wave's coordinates and values, all loads in flight at once
  float c[kOneSteps][3];
  float2 val[kOneSteps];
#pragma unroll
  for (int s = 0; s < kOneSteps; ++s) {
    const int64_t j = j0 + 64 * s;
    const bool valid = j < pa.num_points;
    load_point(pa, dim, j, valid, c[s]);
    val[s] = (REC && pr.values) ? pr.values[valid ? j : 0] : make_float2(0.f, 0.f);
  }
  // fused: this thread's run of tile totals (clamped indices, no branch)
  const int per = (nbins + kOneThreads - 1) / kOneThreads;
  const int t0 = threadIdx.x * per;
  uint32_t tv[kFusedTotPer];
  if (fused) {
#pragma unroll
    for (int i = 0; i < kFusedTotPer; ++i) tv[i] = pr.tot[min(t0 + i, nbins - 1)];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < kFusedTotPer; ++i) sum += (i < per && t0 + i < nbins) ? tv[i] : 0u;
    uint32_t total;
    uint32_t ex = block_scan_waves<kOneThreads>(sum, wsum, total);  // its barriers: rows zeroed
#pragma unroll
    for (int i = 0; i < kFusedTotPer; ++i) {
      if (i < per && t0 + i < nbins) bst[t0 + i] = ex;
      ex += (i < per && t0 + i < nbins) ? tv[i] : 0u;
    }
  } else {
    __syncthreads();  // rows zeroed
  }
  const uint32_t *coff = pr.cnt + (int64_t)chunk * nbins;
  const uint32_t *bin_start = pr.bin_start;
  uint32_t key[kOneSteps], rk[kOneSteps], pos[kOneSteps];
  float zr[kOneSteps][3];
  int off[kOneSteps];
#pragma unroll
  for (int s = 0; s < kOneSteps; ++s) {
    const int64_t j = j0 + 64 * s;
    const bool valid = j < pa.num_points;
    const uint32_t k = valid ? fold_point<REC>(c[s], pr.g, dim, zr[s], off[s]) : 0u;
    // the chunk's offset (and without the fused scan the tile start): loads
    // in flight during the ranking
    pos[s] = coff[k] + (fused ? 0u : bin_start[k]);
    const uint64_t peers = key_peers(k, valid, bits);
    const uint64_t below = peers & lt;
    uint32_t b0 = 0;
    if (valid && below == 0ull) {  // the first lane of each key claims the step's slots
      b0 = row[k];
      row[k] = (uint16_t)(b0 + (uint32_t)__popcll(peers));
    }
    const int leader = valid ? __ffsll((unsigned long long)peers) - 1 : lane;
    b0 = __shfl(b0, leader);
    key[s] = k;
    rk[s] = b0 + (uint32_t)__popcll(below);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < nbins; d += kOneThreads) {
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < kOneWaves; ++w) {
      const uint32_t v = rows[w * nbins + d];
      rows[w * nbins + d] = (uint16_t)run;
      run += v;
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < kOneSteps; ++s) {
    const uint32_t k = key[s];
    pos[s] += (fused ? bst[k] : 0u) + (uint32_t)row[k] + rk[s];
  }
#pragma unroll
  for (int s = 0; s < kOneSteps; ++s) {
    const int64_t j = j0 + 64 * s;
    if (j >= pa.num_points) continue;
    if (REC) {
      pr.out.r[2 * (int64_t)pos[s]] =
          make_float4(zr[s][0], zr[s][1], zr[s][2], __int_as_float(off[s]));
      pr.out.r[2 * (int64_t)pos[s] + 1] =
          make_float4(val[s].x, val[s].y, __uint_as_float(key[s]), __uint_as_float((uint32_t)j));
    } else {
      pr.keys_out[pos[s]] = key[s];
      pr.idx_out[pos[s]] = (uint32_t)j;
    }
  }
}

constexpr int kBinTableChunk = 12;  // descriptors per table-fill launch (kernarg size)
struct BinTableChunk {
  int first, count;
  BinProblem probs[kBinTableChunk];
};
__global__ void bin_table_kernel(BinTableChunk chunk, BinProblem *table) {
  const int i = threadIdx.x;
  if (i < chunk.count) table[chunk.first + i] = chunk.probs[i];
}

int key_bits(int nbins) {
  int bits = 0;
  while ((1 << bits) < nbins) ++bits;
  return bits;
}

hipError_t set_scatter_lds(bool rec, bool batch, size_t lds) {
  if (lds <= 65536) return hipSuccess;
  const void *fn = rec ? (batch ? (const void *)chunk_scatter_kernel<true, true>
                                : (const void *)chunk_scatter_kernel<true, false>)
                       : (batch ? (const void *)chunk_scatter_kernel<false, true>
                                : (const void *)chunk_scatter_kernel<false, false>);
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

} // namespace

bool sort_one_pass(int nbins) { return nbins <= kOneMaxBins; }

size_t sort_count_entries(int64_t num_points, int nbins) {
  const int64_t M = std::max<int64_t>(num_points, 1);
  const int64_t old_n = (int64_t)kRadix * ((M + kSortChunk - 1) / kSortChunk);
  const int64_t one_n = sort_one_pass(nbins) ? (int64_t)nbins * ((M + kOneChunk - 1) / kOneChunk) : 0;
  return (size_t)std::max(old_n, one_n);
}

// One-pass binning: see the comment above chunk_count_kernel.
static hipError_t sort_points_one_pass(const PointArgs &pa, const Geom &g, int dim,
                                       const SortBuffers &sb, hipStream_t s,
                                       const RecordOut *rec) {
  const int64_t n = pa.num_points;
  const int nbins = sb.nbins;
  BinProblem pr{};
  pr.pa = pa;
  pr.g = g;
  pr.cnt = sb.cnt;
  pr.tot = sb.scan_tmp;
  pr.bin_start = sb.bin_start;
  pr.nsub = sb.sub_ofs + sb.nbins;
  pr.subs = sb.subs;
  if (rec) {
    pr.out = rec->out;
    pr.values = rec->values;
  } else {
    pr.keys_out = sb.keys[sb.passes & 1];
    pr.idx_out = sb.idx[sb.passes & 1];
  }
  pr.dim = dim;
  pr.nbins = nbins;
  pr.bits = key_bits(nbins);
  pr.max_sub = sb.max_sub;
  pr.sub_work = sb.sub_work;
  pr.nchunks = (int)((n + kOneChunk - 1) / kOneChunk);
  const int nch = pr.nchunks;
  // with the tile starts in LDS beside the count rows, the scatter launch
  // also writes the list (one more block)
  pr.fused = n > 0 && nbins <= kFusedMaxBins;
  if (n > 0) {
    hipLaunchKernelGGL(chunk_count_kernel<false>, dim3(nch), dim3(kCountThreads),
                       sizeof(uint32_t) * nbins, s, pr, (const BinProblem *)nullptr, 1);
    hipLaunchKernelGGL(chunk_scan_kernel<false>, dim3((nbins + kScanDigits - 1) / kScanDigits),
                       dim3(kScanDigits * kChunkScanSegs), 0, s, pr, (const BinProblem *)nullptr, 1);
  }
  if (!pr.fused) {
    // tile starts and the subproblem list need only the counts
    const int3 nb = dim == 3 ? make_int3(g.nbin[0], g.nbin[1], g.nbin[2]) : make_int3(1, 1, 1);
    hipLaunchKernelGGL(subproblems_kernel, dim3(1), dim3(kSubThreads), 0, s, sb.bin_start,
                       n > 0 ? (const uint32_t *)sb.scan_tmp : (const uint32_t *)nullptr, 1,
                       sb.nbins, nb, sb.max_sub, sb.sub_work, sb.sub_ofs + sb.nbins, sb.subs);
  }
  if (n > 0) {
    size_t lds = sizeof(uint16_t) * (size_t)kOneWaves * nbins;
    if (pr.fused) lds = std::max(lds + sizeof(uint32_t) * nbins, sizeof(SubLds<kOneThreads>));
    hipError_t e = set_scatter_lds(rec != nullptr, false, lds);
    if (e != hipSuccess) return e;
    const int blocks = nch + (pr.fused ? 1 : 0);
    if (rec)
      hipLaunchKernelGGL((chunk_scatter_kernel<true, false>), dim3(blocks), dim3(kOneThreads), lds,
                         s, pr, (const BinProblem *)nullptr, 1);
    else
      hipLaunchKernelGGL((chunk_scatter_kernel<false, false>), dim3(blocks), dim3(kOneThreads),
                         lds, s, pr, (const BinProblem *)nullptr, 1);
  }
  return hipGetLastError();
}

hipError_t sort_points_batched(BinProblem *probs, int nprob, BinProblem *table, hipStream_t s) {
  if (nprob <= 0) return hipSuccess;
  uint32_t chunks = 0, scans = 0;
  int max_bins = 1;
  bool rec = probs[0].out.r != nullptr;
  for (int i = 0; i < nprob; ++i) {
    BinProblem &pr = probs[i];
    if (!sort_one_pass(pr.nbins) || (pr.out.r != nullptr) != rec) return hipErrorInvalidValue;
    pr.bits = key_bits(pr.nbins);
    pr.nchunks = (int)((pr.pa.num_points + kOneChunk - 1) / kOneChunk);
    pr.chunk0 = chunks;
    pr.scan0 = scans;
    chunks += (uint32_t)pr.nchunks;
    if (pr.pa.num_points > 0) scans += (uint32_t)((pr.nbins + kScanDigits - 1) / kScanDigits);
    max_bins = std::max(max_bins, pr.nbins);
  }
  for (int first = 0; first < nprob; first += kBinTableChunk) {
    BinTableChunk chunk{};
    chunk.first = first;
    chunk.count = std::min(kBinTableChunk, nprob - first);
    for (int i = 0; i < chunk.count; ++i) chunk.probs[i] = probs[first + i];
    hipLaunchKernelGGL(bin_table_kernel, dim3(1), dim3(64), 0, s, chunk, table);
  }
  const BinProblem none{};
  if (chunks > 0) {
    hipLaunchKernelGGL(chunk_count_kernel<true>, dim3(chunks), dim3(kCountThreads),
                       sizeof(uint32_t) * max_bins, s, none, (const BinProblem *)table, nprob);
    if (scans > 0)
      hipLaunchKernelGGL(chunk_scan_kernel<true>, dim3(scans), dim3(kScanDigits * kChunkScanSegs),
                         0, s, none, (const BinProblem *)table, nprob);
  }
  hipLaunchKernelGGL(subproblems_batched_kernel, dim3(nprob), dim3(kSubThreads), 0, s,
                     (const BinProblem *)table);
  if (chunks > 0) {
    const size_t lds = sizeof(uint16_t) * (size_t)kOneWaves * max_bins;
    hipError_t e = set_scatter_lds(rec, true, lds);
    if (e != hipSuccess) return e;
    if (rec)
      hipLaunchKernelGGL((chunk_scatter_kernel<true, true>), dim3(chunks), dim3(kOneThreads), lds,
                         s, none, (const BinProblem *)table, nprob);
    else
      hipLaunchKernelGGL((chunk_scatter_kernel<false, true>), dim3(chunks), dim3(kOneThreads), lds,
                         s, none, (const BinProblem *)table, nprob);
  }
  return hipGetLastError();
}

hipError_t sort_points(const PointArgs &pa, const Geom &g, int dim, const SortBuffers &sb,
                       hipStream_t s, const RecordOut *rec) {
  if (sort_one_pass(sb.nbins)) return sort_points_one_pass(pa, g, dim, sb, s, rec);
  const int64_t n = pa.num_points;
  // with records, the last pass writes rec->out; earlier ones alternate with tmp
  auto dst = [&](int q) { return ((sb.passes - 1 - q) & 1) == 0 ? rec->out : rec->tmp; };
  if (n > 0) {
    hipLaunchKernelGGL(keys_kernel, dim3(sb.nchunks), dim3(kSortThreads), 0, s, pa, g, dim,
                       sb.keys[0], sb.cnt, sb.nchunks);
    for (int p = 0; p < sb.passes; ++p) {
      if (p > 0) {
        const KeySrc ks{sb.keys[p & 1], rec ? dst(p - 1).r : nullptr};
        hipLaunchKernelGGL(hist_kernel, dim3(sb.nchunks), dim3(kSortThreads), 0, s, ks, n,
                           p * kRadixBits, sb.cnt, sb.nchunks);
      }
      hipLaunchKernelGGL(scan_rows_kernel, dim3(kRadix), dim3(kScanThreads), 0, s, sb.cnt,
                         sb.nchunks, sb.scan_tmp);
      RecMove rm{};
      if (rec) {
        rm.out = dst(p);
        if (p > 0) rm.in = dst(p - 1);
        rm.values = rec->values;
        rm.pa = pa;
        rm.g = g;
        rm.first = p == 0;
        hipLaunchKernelGGL(scatter_kernel<true>, dim3(sb.nchunks), dim3(kSortThreads), 0, s,
                           sb.keys[0], (const uint32_t *)nullptr, (uint32_t *)nullptr,
                           (uint32_t *)nullptr, n, p * kRadixBits, sb.cnt, sb.scan_tmp,
                           sb.nchunks, rm);
      } else {
        hipLaunchKernelGGL(scatter_kernel<false>, dim3(sb.nchunks), dim3(kSortThreads), 0, s,
                           sb.keys[p & 1], p == 0 ? (const uint32_t *)nullptr : sb.idx[p & 1],
                           sb.keys[(p + 1) & 1], sb.idx[(p + 1) & 1], n, p * kRadixBits, sb.cnt,
                           sb.scan_tmp, sb.nchunks, rm);
      }
    }
  }
  const KeySrc sorted{sb.keys[sb.passes & 1], rec ? rec->out.r : nullptr};
  int nb1 = sb.nbins + 1;
  const int3 nb = dim == 3 ? make_int3(g.nbin[0], g.nbin[1], g.nbin[2]) : make_int3(1, 1, 1);
  hipLaunchKernelGGL(bounds_kernel, dim3((nb1 + 255) / 256), dim3(256), 0, s, sorted, n,
                     sb.bin_start, sb.nbins);
  hipLaunchKernelGGL(subproblems_kernel, dim3(1), dim3(kSubThreads), 0, s, sb.bin_start,
                     (const uint32_t *)nullptr, 0, sb.nbins, nb, sb.max_sub, sb.sub_work,
                     sb.sub_ofs + sb.nbins, sb.subs);
  return hipGetLastError();
}

} // namespace anufft
