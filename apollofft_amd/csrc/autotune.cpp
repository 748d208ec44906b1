// autotune.cpp -- the autotuner of the launch settings of a 3D plan
// (anufft_autotune3d, declared in include/apollo_nufft.h).
//
// The reference fixes its binning by a rule (bucket_count: ceil(m / (2w+1))
// buckets per axis, apollo-nufft/src/infrastructure/kernel/kaiser_bessel.rs:
// 162-164).  Here the tile geometry follows the kernels, and what is left
// free -- points per subproblem, the work weight that splits the subproblem
// list among workgroups, workgroups per CU of the point kernels, FFT lines
// per workgroup and the kernel method -- is chosen by timing: coordinate
// descent over candidate values, each candidate a plan built with those
// options and timed (HIP events, best of three after one warm-up) on the
// caller's points or on uniform synthetic points.  The winner can be kept in
// a JSON cache keyed by device, grid, kernel width, point count (rounded up
// to a power of two), workload and point source.  The tuner uses only the
// public C ABI; it allocates device memory (it is not meant for hot loops).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/apollo_nufft.h"

namespace anufft {
anufft_status report_error(anufft_status st, const char *msg);
}

namespace {

struct Knobs {
  int kernel_method, max_subproblem, sub_work, blocks_per_cu, fft_lines;
};

const char *kFields[] = {"kernel_method", "max_subproblem", "sub_work", "blocks_per_cu",
                         "fft_lines"};

int &knob(Knobs &k, int i) {
  switch (i) {
  case 0: return k.kernel_method;
  case 1: return k.max_subproblem;
  case 2: return k.sub_work;
  case 3: return k.blocks_per_cu;
  default: return k.fft_lines;
  }
}

int knob(const Knobs &k, int i) { return knob(const_cast<Knobs &>(k), i); }

void apply(anufft_opts &o, const Knobs &k) {
  o.kernel_method = k.kernel_method;
  o.max_subproblem = k.max_subproblem;
  o.sub_work = k.sub_work;
  o.blocks_per_cu = k.blocks_per_cu;
  o.fft_lines = k.fft_lines;
}

// ---- JSON cache: {"<key>": {"field": value, ...}, ...}, one entry per line --
std::mutex g_cache_mu;

bool parse_field(const std::string &line, const char *name, double &v) {
  const std::string pat = std::string("\"") + name + "\":";
  size_t p = line.find(pat);
  if (p == std::string::npos) return false;
  p += pat.size();
  char *end = nullptr;
  v = std::strtod(line.c_str() + p, &end);
  return end != line.c_str() + p;
}

std::map<std::string, std::string> read_cache(const char *path) {
  std::map<std::string, std::string> entries;
  std::ifstream in(path);
  std::string line;
  while (std::getline(in, line)) {
    const size_t a = line.find('"');
    const size_t b = a == std::string::npos ? a : line.find("\": {", a + 1);
    if (a == std::string::npos || b == std::string::npos) continue;
    std::string v = line.substr(b + 3);
    while (!v.empty() && (v.back() == ',' || v.back() == ' ' || v.back() == '\r')) v.pop_back();
    entries[line.substr(a + 1, b - a - 1)] = v;
  }
  return entries;
}

bool cache_lookup(const char *path, const std::string &key, Knobs &k, double &ms) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  auto entries = read_cache(path);
  auto it = entries.find(key);
  if (it == entries.end()) return false;
  Knobs r{};
  for (int i = 0; i < 5; ++i) {
    double v;
    if (!parse_field(it->second, kFields[i], v)) return false;
    knob(r, i) = (int)v;
  }
  if (!parse_field(it->second, "ms", ms)) return false;
  k = r;
  return true;
}

bool cache_store(const char *path, const std::string &key, const Knobs &k, double ms) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  auto entries = read_cache(path);
  std::ostringstream e;
  e << "{";
  for (int i = 0; i < 5; ++i) e << "\"" << kFields[i] << "\": " << knob(k, i) << ", ";
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.6f", ms);
  e << "\"ms\": " << buf << "}";
  entries[key] = e.str();
  const std::string tmp = std::string(path) + ".tmp";
  {
    std::ofstream out(tmp, std::ios::trunc);
    if (!out) return false;
    out << "{\n";
    size_t i = 0;
    for (const auto &kv : entries)
      out << "\"" << kv.first << "\": " << kv.second << (++i < entries.size() ? ",\n" : "\n");
    out << "}\n";
    if (!out) return false;
  }
  return std::rename(tmp.c_str(), path) == 0;
}

// ---- timing -------------------------------------------------------------------
struct Bench {
  int64_t n[3];
  double d[3];
  int sigma, w;
  anufft_opts base;
  int64_t M;
  const float *coords[3];
  int64_t stride;
  int workload;
  const float *values = nullptr, *modes = nullptr;
  float *out1 = nullptr, *out2 = nullptr;
  void *ws = nullptr;
  size_t ws_bytes = 0;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;

  // Best of three timed runs (after one warm-up) of the workload with the
  // plan built from k; +inf when the plan or a run fails.
  double time(const Knobs &k) {
    anufft_opts o = base;
    apply(o, k);
    anufft_plan *p = nullptr;
    if (anufft_plan3d_create(n[0], n[1], n[2], d[0], d[1], d[2], sigma, w, &o, &p) != ANUFFT_OK)
      return INFINITY;
    double best = INFINITY;
    size_t need = 0;
    if (anufft_workspace_size(p, M, &need) == ANUFFT_OK) {
      bool ok = true;
      if (need > ws_bytes) {
        if (ws) (void)hipFree(ws);
        ws = nullptr;
        ws_bytes = 0;
        ok = hipMalloc(&ws, need) == hipSuccess;
        if (ok) ws_bytes = need;
      }
      for (int r = 0; ok && r < 4; ++r) {
        ok = hipEventRecord(e0, s) == hipSuccess;
        if (ok && (workload & 1))
          ok = anufft_execute_type1(p, M, coords, stride, values, out1, ws, ws_bytes, s) == ANUFFT_OK;
        if (ok && (workload & 2))
          ok = anufft_execute_type2(p, M, coords, stride, modes, out2, ws, ws_bytes, s) == ANUFFT_OK;
        ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
        float ms = 0.f;
        if (ok && r > 0 && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) best = std::min(best, (double)ms);
      }
    }
    anufft_plan_destroy(p);
    return best;
  }
};

std::string device_tag(int dev) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
  std::ostringstream t;
  t << prop.gcnArchName << "/" << prop.multiProcessorCount << "cu";
  std::string s = t.str();
  for (char &c : s)
    if (c == '"' || c == ':') c = '_';
  return s;
}

} // namespace

extern "C" anufft_status anufft_autotune3d(int64_t nx, int64_t ny, int64_t nz, double dx,
                                           double dy, double dz, int32_t sigma,
                                           int32_t kernel_width, const anufft_opts *base,
                                           int64_t num_points, const float *const coords[3],
                                           int64_t coord_stride, int32_t workload,
                                           const char *cache_path, anufft_opts *tuned,
                                           double *best_ms) {
  using anufft::report_error;
  if (!tuned) return report_error(ANUFFT_INVALID_ARG, "tuned options out-pointer is NULL");
  if (num_points < 1 || num_points > 2147483647)
    return report_error(ANUFFT_INVALID_ARG, "autotune needs 1 .. 2^31 - 1 points");
  if (workload < 1 || workload > 3)
    return report_error(ANUFFT_INVALID_ARG, "workload must be 1 (Type-1), 2 (Type-2) or 3 (both)");
  if (coords && (!coords[0] || !coords[1] || !coords[2] || coord_stride < 1))
    return report_error(ANUFFT_INVALID_ARG, "coordinate pointers or stride invalid");
  anufft_opts b;
  if (base) b = *base;
  else anufft_default_opts(&b);
  // the geometry must make a valid plan (validation errors come from here)
  anufft_plan *probe = nullptr;
  anufft_status st = anufft_plan3d_create(nx, ny, nz, dx, dy, dz, sigma, kernel_width, &b, &probe);
  if (st != ANUFFT_OK) return st;
  anufft_meta meta;
  st = anufft_plan_get_meta(probe, &meta);
  anufft_plan_destroy(probe);
  if (st != ANUFFT_OK) return st;
  int dev = b.device_id;
  if (dev < 0 && hipGetDevice(&dev) != hipSuccess)
    return report_error(ANUFFT_DEVICE_UNAVAILABLE, "no HIP device available");

  // start: the automatic settings (the method the base options ask for)
  Knobs cur{b.kernel_method == 0 ? meta.kernel_method : b.kernel_method, meta.max_subproblem,
            meta.sub_work, 0, 0};
  int64_t mb = 1;
  while (mb < num_points) mb <<= 1;
  std::ostringstream key;
  key << device_tag(dev) << "|n=" << nx << "," << ny << "," << nz << "|m=" << meta.m[0] << ","
      << meta.m[1] << "," << meta.m[2] << "|w=" << kernel_width << "|sigma=" << sigma
      << "|method=" << b.kernel_method << "|M=" << mb << "|workload=" << workload
      << "|points=" << (coords ? "caller" : "uniform");
  double ms = 0.0;
  if (cache_path && cache_lookup(cache_path, key.str(), cur, ms)) {
    *tuned = b;
    apply(*tuned, cur);
    if (best_ms) *best_ms = ms;
    return ANUFFT_OK;
  }

  // device buffers: coordinates (caller's or uniform on [-L/2, L/2)), values, modes, outputs
  int prev = -1;
  (void)hipGetDevice(&prev);
  if (prev != dev) (void)hipSetDevice(dev);
  Bench bench{};
  bench.n[0] = nx; bench.n[1] = ny; bench.n[2] = nz;
  bench.d[0] = dx; bench.d[1] = dy; bench.d[2] = dz;
  bench.sigma = sigma;
  bench.w = kernel_width;
  bench.base = b;
  bench.M = num_points;
  bench.workload = workload;
  std::vector<void *> owned;
  auto dalloc = [&](size_t bytes) -> void * {
    void *q = nullptr;
    if (hipMalloc(&q, bytes) != hipSuccess) return nullptr;
    owned.push_back(q);
    return q;
  };
  const size_t M = (size_t)num_points, N = (size_t)(nx * ny * nz);
  std::mt19937_64 rng(20240917u);
  std::normal_distribution<float> nd(0.f, 1.f);
  bool ok = hipStreamCreate(&bench.s) == hipSuccess && hipEventCreate(&bench.e0) == hipSuccess &&
            hipEventCreate(&bench.e1) == hipSuccess;
  if (ok && coords) {
    for (int a = 0; a < 3; ++a) bench.coords[a] = coords[a];
    bench.stride = coord_stride;
  } else if (ok) {
    std::vector<float> h(3 * M);
    std::uniform_real_distribution<double> ud(-0.5, 0.5);
    const double L[3] = {nx * dx, ny * dy, nz * dz};
    for (size_t j = 0; j < M; ++j)
      for (int a = 0; a < 3; ++a) h[3 * j + a] = (float)(ud(rng) * L[a]);
    float *dc = (float *)dalloc(sizeof(float) * 3 * M);
    ok = dc && hipMemcpy(dc, h.data(), sizeof(float) * 3 * M, hipMemcpyHostToDevice) == hipSuccess;
    for (int a = 0; a < 3; ++a) bench.coords[a] = dc ? dc + a : nullptr;
    bench.stride = 3;
  }
  if (ok) {
    std::vector<float> h(2 * std::max(M, N));
    for (float &v : h) v = nd(rng);
    float *dv = (float *)dalloc(sizeof(float) * 2 * M);
    float *dm = (float *)dalloc(sizeof(float) * 2 * N);
    bench.out1 = (float *)dalloc(sizeof(float) * 2 * N);
    bench.out2 = (float *)dalloc(sizeof(float) * 2 * M);
    ok = dv && dm && bench.out1 && bench.out2 &&
         hipMemcpy(dv, h.data(), sizeof(float) * 2 * M, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(dm, h.data(), sizeof(float) * 2 * N, hipMemcpyHostToDevice) == hipSuccess;
    bench.values = dv;
    bench.modes = dm;
  }

  // coordinate descent over the candidate values of each setting
  double best = ok ? bench.time(cur) : INFINITY;
  if (ok && std::isfinite(best)) {
    std::vector<std::vector<int>> cand(5);
    if (b.kernel_method == 0) cand[0] = {2, 1};
    cand[1] = {1024, 2048, 4096, 8192, 16384};
    cand[2] = {8, 16, 48, 128, 384};
    cand[3] = {0, 1, 2, 3, 4};
    for (int lb : {0, 1, 2, 4, 8, 16, 32})
      if (lb * (int)std::max(meta.m[0], std::max(meta.m[1], meta.m[2])) <= 4096) cand[4].push_back(lb);
    for (int i = 0; i < 5; ++i)
      for (int v : cand[i]) {
        if (v == knob(cur, i)) continue;
        Knobs k = cur;
        knob(k, i) = v;
        const double t = bench.time(k);
        if (t < best) {
          best = t;
          cur = k;
        }
      }
  }
  for (void *q : owned) (void)hipFree(q);
  if (bench.ws) (void)hipFree(bench.ws);
  if (bench.e0) (void)hipEventDestroy(bench.e0);
  if (bench.e1) (void)hipEventDestroy(bench.e1);
  if (bench.s) (void)hipStreamDestroy(bench.s);
  if (prev >= 0 && prev != dev) (void)hipSetDevice(prev);
  if (!ok || !std::isfinite(best))
    return report_error(ANUFFT_HIP_ERROR, "autotune: device buffers or timed runs failed");
  if (cache_path && !cache_store(cache_path, key.str(), cur, best))
    return report_error(ANUFFT_INVALID_ARG, "autotune: cannot write the cache file");
  *tuned = b;
  apply(*tuned, cur);
  if (best_ms) *best_ms = best;
  return ANUFFT_OK;
}
