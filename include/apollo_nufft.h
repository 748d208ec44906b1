/*
 * apollo_nufft.h -- C ABI of the MI355X (gfx950) Kaiser-Bessel NUFFT library
 * (libapollo_nufft_hip.so).
 *
 * Drop-in boundary for apollo-nufft's plan/execute API.  Each entry point
 * names the reference interface it replaces (paths relative to the apollofft
 * repository root).  Conventions kept from the reference:
 *   - kernel_width is the KB half-width w: 2w+1 taps per axis, 2w non-zero
 *     (apollo-nufft/src/infrastructure/kernel/kaiser_bessel.rs:109-116);
 *   - beta = pi (1 - 1/(2 sigma)) 2w with the integer sigma
 *     (apollo-nufft/src/application/execution/transform/dimension_3d.rs:200);
 *   - Type-1 f_k = sum_j c_j exp(-2 pi i k.x_j/L), Type-2
 *     g_j = sum_k f_k exp(+2 pi i k.x_j/L) (dimension_1d.rs:10-19);
 *   - modes in FFT order with k <= n/2 taken as positive
 *     (kaiser_bessel.rs:152-158), 3D arrays row-major [kx][ky][kz]
 *     (dimension_3d.rs:400-413, ndarray Array3);
 *   - positions may be anywhere on the real line; they are folded into
 *     [0, L) per axis (dimension_3d.rs:93-95).
 * Complex data are interleaved float pairs (Complex32 / torch.complex64).
 *
 * Ownership: plans own read-only device tables (deconvolution factors in
 * FP64, FP64-built twiddles, kernel polynomial coefficients).  The caller
 * owns inputs, outputs and the workspace; execute calls allocate nothing.
 * Execution is asynchronous on the given hipStream_t (NULL = legacy default
 * stream).  Plans are immutable after creation and may be shared between
 * threads; calls that share one workspace must not overlap.
 * Every call returns a status; anufft_last_error() gives the message of the
 * last failure on the calling thread.  Nothing aborts.
 */
#ifndef APOLLO_NUFFT_H
#define APOLLO_NUFFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Mirrors apollo-fft ApolloError (apollo-fft/src/domain/
 * contracts/error.rs:10-63), NufftMetadataError (apollo-nufft/src/domain/
 * metadata/grid.rs:7-23) and NufftWgpuError (apollo-nufft-wgpu/src/domain/
 * error.rs:10-49). */
typedef enum anufft_status {
  ANUFFT_OK = 0,
  ANUFFT_INVALID_ARG = 1,        /* NufftMetadataError, sigma/w < 2, InvalidPlan */
  ANUFFT_SHAPE_MISMATCH = 2,     /* ShapeMismatch / InputLengthMismatch */
  ANUFFT_PROFILE_MISMATCH = 3,   /* Validation{field:"precision_profile"} */
  ANUFFT_CAPACITY_EXCEEDED = 4,  /* workspace smaller than anufft_workspace_size */
  ANUFFT_U32_OVERFLOW = 5,       /* validate_usize_to_u32 (device.rs:973-980) */
  ANUFFT_DEVICE_UNAVAILABLE = 6, /* AdapterUnavailable / DeviceUnavailable */
  ANUFFT_HIP_ERROR = 7,          /* HIP runtime error */
  ANUFFT_UNSUPPORTED = 8         /* UnsupportedExecution */
} anufft_status;

/* Oversampled-grid policy. */
typedef enum anufft_grid_policy {
  /* Reference rule: 3D m = next_pow2(max(sigma n, 2w+1)) (dimension_3d.rs:48-54,
   * device.rs:856-891); 1D m = sigma n (dimension_1d.rs:131). */
  ANUFFT_GRID_REFERENCE = 0,
  /* m = max(sigma n, 2w+1) exactly: mixed-radix grids such as 200 or 254. */
  ANUFFT_GRID_EXACT = 1
} anufft_grid_policy;

typedef struct anufft_opts {
  int32_t grid_policy;                       /* anufft_grid_policy */
  int32_t compat_reference_type2_3d_scaling; /* 1: 3D T2 returns direct/(mx my mz) like the reference */
  int32_t device_id;                         /* -1: current device */
  int32_t bin_size[3];                       /* interpolation tile in cells; 0 = automatic */
  int32_t max_subproblem;                    /* points per subproblem; 0 = automatic */
  int32_t kernel_method;                     /* 3D spreading and interpolation kernels:
                                                0 automatic, 1 vector ALU (register columns,
                                                LDS tiles), 2 FP16-split MFMA blocks */
  /* Launch settings that change speed, never results (0 = automatic; the
   * values anufft_autotune3d picks): */
  int32_t sub_work;      /* fixed cost of a subproblem, in units of 16 points, when the
                            subproblem list is split among workgroups (automatic: 48) */
  int32_t blocks_per_cu; /* workgroups per CU of the 3D point kernels (automatic: per kernel) */
  int32_t fft_lines;     /* lines per workgroup of the FFT passes (automatic: 1024 / length,
                            at least 2 workgroups per CU) */
} anufft_opts;

typedef struct anufft_meta {
  int32_t dim;          /* 1 or 3 */
  int64_t n[3];         /* modes per axis (1D: n[1] = n[2] = 1) */
  int64_t m[3];         /* oversampled grid per axis */
  double length[3];     /* periodic lengths L = n dx */
  int32_t kernel_width; /* w (half-width) */
  int32_t sigma;
  double beta;
  double i0_beta;
  int32_t grid_policy;
  int32_t compat_reference_type2_3d_scaling;
  int32_t bin_size[3];        /* spreading tile (cells) */
  int32_t max_subproblem;     /* points per spreading subproblem */
  int32_t interp_bin_size[3]; /* interpolation / gradient tile (cells) */
  int32_t interp_max_subproblem;
  int32_t poly_degree;  /* degree of the per-tap kernel polynomials */
  double poly_max_err;  /* max |phi - fit| on [-w, w] measured at plan creation */
  int32_t kernel_method; /* 3D spread/interp kernels in use (1 vector ALU, 2 MFMA) */
  int32_t sub_work;      /* launch settings in use (see anufft_opts) */
  int32_t blocks_per_cu; /* 0: automatic */
  int32_t fft_lines;     /* 0: automatic */
} anufft_meta;

typedef struct anufft_plan anufft_plan;

/* Fill opts with defaults (reference grid policy, automatic tiles). */
void anufft_default_opts(anufft_opts *opts);

/* NufftPlan1D::new (apollo-nufft .../transform/dimension_1d.rs:127-154),
 * NufftWgpuPlan1D::new (apollo-nufft-wgpu/src/application/plan.rs:16) and
 * UniformDomain1D::new validation (domain/metadata/grid.rs:36-48). */
anufft_status anufft_plan1d_create(int64_t n, double dx, int32_t sigma, int32_t kernel_width,
                                   const anufft_opts *opts, anufft_plan **plan);

/* NufftPlan3D::new (dimension_3d.rs:193-232), NufftWgpuPlan3D::new
 * (application/plan.rs:54), UniformGrid3D::new validation (grid.rs:76-110). */
anufft_status anufft_plan3d_create(int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                                   double dz, int32_t sigma, int32_t kernel_width,
                                   const anufft_opts *opts, anufft_plan **plan);

anufft_status anufft_plan_destroy(anufft_plan *plan);

/* KB metadata of the plan: the private fields m, beta, i0_beta of
 * NufftPlan{1D,3D} (dimension_1d.rs:102-111, dimension_3d.rs:164-178) and the
 * GPU copy in fast_{1d,3d}_metadata (apollo-nufft-wgpu .../device.rs:842-1034). */
anufft_status anufft_plan_get_meta(const anufft_plan *plan, anufft_meta *meta);

/* FP64 deconvolution factors 1/phi_hat(k_signed/m) of one axis, n values
 * (axis_deconv, kaiser_bessel.rs:137-148). */
anufft_status anufft_plan_get_deconv(const anufft_plan *plan, int32_t axis, double *out_n);

/* Workspace bytes for executing up to max_points points with this plan
 * (replaces NufftGpuBuffers{1D,3D}::new capacity, apollo-nufft-wgpu
 * .../kernel.rs:82-181,234). */
anufft_status anufft_workspace_size(const anufft_plan *plan, int64_t max_points, size_t *bytes);

/* Type-1, device pointers.  Replaces NufftPlan{1D,3D}::type1_into
 * (dimension_1d.rs:166, dimension_3d.rs:256) and NufftWgpuBackend::
 * execute_fast_type1_{1d,3d}[_with_buffers] (device.rs:217,383,681,752).
 * Point coordinate j of axis a is read at coords[a][j * coord_stride]
 * (SoA: stride 1; AoS (x,y,z) tuples: stride 3; a stride below 1 is
 * ANUFFT_INVALID_ARG).  1D plans read coords[0] only.  values: M complex;
 * out: n (1D) or nx*ny*nz (3D) complex.  Every execute call takes at most
 * 2^31 - 1 points (more: ANUFFT_U32_OVERFLOW). */
anufft_status anufft_execute_type1(const anufft_plan *plan, int64_t num_points,
                                   const float *const coords[3], int64_t coord_stride,
                                   const float *values, float *out, void *workspace,
                                   size_t workspace_bytes, void *hip_stream);

/* Type-1 without a workspace argument and without any allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and stream
 * semantics as anufft_execute_type1; the workspace is a fixed scratch array
 * of anufft_scratch_capacity() bytes (64 MiB) in device memory, one per GPU,
 * held until the process ends.  Calls needing more than that
 * (anufft_workspace_size(plan, num_points) > capacity; e.g. about 500k points
 * on a 128^3 grid, about 800k on a 64^3 grid) return
 * ANUFFT_CAPACITY_EXCEEDED.  Every call on one GPU shares that GPU's scratch:
 * calls that overlap in time on one GPU (other streams, other threads) race
 * and must be serialized by the caller. */
anufft_status anufft_execute_type1_noalloc(const anufft_plan *plan, int64_t num_points,
                                           const float *const coords[3], int64_t coord_stride,
                                           const float *values, float *out, void *hip_stream);
/* Size in bytes of the scratch array behind anufft_execute_type1_noalloc. */
anufft_status anufft_scratch_capacity(size_t *bytes);

/* Type-2, device pointers.  Replaces NufftPlan{1D,3D}::type2_into
 * (dimension_1d.rs:224, dimension_3d.rs:496) and execute_fast_type2_{1d,3d}
 * (device.rs:277,446).  modes: n or nx*ny*nz complex; out: M complex. */
anufft_status anufft_execute_type2(const anufft_plan *plan, int64_t num_points,
                                   const float *const coords[3], int64_t coord_stride,
                                   const float *modes, float *out, void *workspace,
                                   size_t workspace_bytes, void *hip_stream);

/* Gradient of a real loss with respect to the point coordinates, for both
 * transform types (no reference counterpart; the reference has no autograd).
 * Given per-point complex u_j and modes h, writes for each axis a
 *   grad[a][j] = Re( conj(u_j) * d/dx_j^(a) [ sum_k h_k exp(+2 pi i k.x_j/L) ] )
 * using the kernel derivative on the oversampled grid.
 *   Type-2 forward c = T2(f,x), upstream gradient g on c:  u = g, h = f.
 *   Type-1 forward F = T1(c,x), upstream gradient G on F:  u = c, h = G.
 * grad[a] may be NULL to skip an axis (3D) and must be non-NULL for a=0. */
anufft_status anufft_execute_grad_points(const anufft_plan *plan, int64_t num_points,
                                         const float *const coords[3], int64_t coord_stride,
                                         const float *u, const float *modes_h,
                                         float *const grad[3], void *workspace,
                                         size_t workspace_bytes, void *hip_stream);

/* Backward of a 3D Type-2 c = T2(f, x) for an upstream gradient g on c (no
 * reference counterpart; the reference has no autograd): writes
 *   grad_modes = T1(g, x)                                  (mode shape)
 *   grad_x[a]  = anufft_execute_grad_points(x, u = g, h = f)[a]
 * from ONE binning of the points carrying g (the Type-1 and the gradient
 * then share the records), with the fine grid of f built on the plan's
 * second stream during the binning.  Same results as the two separate
 * calls.  3D plans whose gradient runs on the MFMA kernels (kernel_method
 * 0 or 2, w <= 7); other plans return ANUFFT_UNSUPPORTED.  grad_modes must
 * be non-NULL, grad_x[0] too when num_points > 0 (grad_x[1], grad_x[2] may be
 * NULL).  The workspace is twice anufft_workspace_size(plan, num_points)
 * bytes. */
anufft_status anufft_execute_type2_backward(const anufft_plan *plan, int64_t num_points,
                                            const float *const coords[3], int64_t coord_stride,
                                            const float *g, const float *modes,
                                            float *grad_modes, float *const grad_x[3],
                                            void *workspace, size_t workspace_bytes,
                                            void *hip_stream);

/* Backward of a 3D Type-1 F = T1(c, x) for an upstream gradient G on F (no
 * reference counterpart; the reference has no autograd): writes
 *   grad_values = T2(G, x)                 (M complex, direct-sum scaling)
 *   grad_x[a]   = anufft_execute_grad_points(x, u = c, h = G)[a]
 * from one FFT of G and one binning of the points carrying c.  Same results
 * as the two separate calls (with the direct-sum scaling).  3D plans whose
 * gradient runs on the MFMA kernels (kernel_method 0 or 2, w <= 7); other
 * plans return ANUFFT_UNSUPPORTED.  grad_values and grad_x[0] must be
 * non-NULL when num_points > 0.  Workspace: anufft_workspace_size bytes. */
anufft_status anufft_execute_type1_backward(const anufft_plan *plan, int64_t num_points,
                                            const float *const coords[3], int64_t coord_stride,
                                            const float *values, const float *G,
                                            float *grad_values, float *const grad_x[3],
                                            void *workspace, size_t workspace_bytes,
                                            void *hip_stream);

/* Batched Type-1 over independent 3D plans (no reference counterpart: the
 * reference has no batched NUFFT).  Problem p uses plans[p], num_points[p],
 * coords_p[a] = coords[3p + a] (stride coord_stride), values[p], out[p].  The
 * workspace must hold anufft_batch_workspace_size bytes.  Work is queued on
 * `hip_stream`; parts of it are forked to the second stream of plans[0] and
 * joined back.  When all plans have the same kernel width, sigma and kernel
 * method, the problems share the binning launches (one count, scan, list and
 * scatter launch for all, when every plan has at most 9216 tiles) and one
 * persistent spreading launch; the FFT passes run per problem, split
 * between the caller's stream and plans[0]'s second stream.  Other mixes
 * run one anufft_execute_type1 per problem. */
anufft_status anufft_batch_workspace_size(const anufft_plan *const *plans, int32_t num_problems,
                                          const int64_t *max_points, size_t *bytes);
anufft_status anufft_execute_type1_batched(const anufft_plan *const *plans, int32_t num_problems,
                                           const int64_t *num_points,
                                           const float *const *coords, int64_t coord_stride,
                                           const float *const *values, float *const *out,
                                           void *workspace, size_t workspace_bytes,
                                           void *hip_stream);

/* Autotuner of the launch settings of a 3D plan geometry.  No counterpart
 * in the reference, whose binning follows a fixed rule (bucket_count,
 * apollo-nufft/src/infrastructure/kernel/kaiser_bessel.rs:162-164).  Times
 * candidate values of kernel_method (when base->kernel_method is 0),
 * max_subproblem, sub_work, blocks_per_cu and fft_lines by coordinate
 * descent, each candidate a plan built with them and timed with HIP events
 * (best of three runs) on `workload` (1 Type-1, 2 Type-2, 3 both) over
 * num_points points: the caller's device coordinates (coords, stride as
 * for execute) or, with coords NULL, uniform synthetic points.  *tuned =
 * *base (defaults when NULL) with the fastest settings; *best_ms = its
 * time per workload.  With cache_path the result is first looked up in and
 * then stored to a JSON file keyed by device, grid, kernel width, sigma,
 * method, point count (rounded up to a power of two), workload and point
 * source.  Allocates device memory and synchronizes (not for hot loops);
 * results of the tuned plan equal those of any other settings up to
 * floating-point summation order. */
anufft_status anufft_autotune3d(int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                                double dz, int32_t sigma, int32_t kernel_width,
                                const anufft_opts *base, int64_t num_points,
                                const float *const coords[3], int64_t coord_stride,
                                int32_t workload, const char *cache_path, anufft_opts *tuned,
                                double *best_ms);

/* Host-pointer variants for drop-in parity tests: copy in, execute on an
 * internal stream of the plan's device, copy out, synchronize.  They allocate
 * device memory per call (not for hot loops).  Coordinates are FP64 host
 * arrays (like NufftPlan{1D,3D}::type1_into, dimension_1d.rs:166) rounded
 * once to FP32 for the device, like the WGPU backend's f32 positions
 * (device.rs:383).  Values use `storage`; the typed reference paths
 * (type1_typed_into / type2_typed_into, dimension_1d.rs:283-349,
 * dimension_3d.rs:333-389,605-629, WGPU device.rs:142-675) require `profile`
 * to match the storage exactly (validate_profile, dimension_1d.rs:403-419),
 * else ANUFFT_PROFILE_MISMATCH.  Arithmetic is FP32 in every case (the WGPU
 * backend's rule, device.rs:248-252); FP16 storage is promoted once on input
 * and quantized once on output (dimension_1d.rs:388-401). */
typedef enum anufft_storage {
  ANUFFT_STORAGE_C128 = 0,  /* Complex64 (f64 pairs)  <-> HIGH_ACCURACY_F64 */
  ANUFFT_STORAGE_C64 = 1,   /* Complex32 (f32 pairs)  <-> LOW_PRECISION_F32 */
  ANUFFT_STORAGE_F16X2 = 2  /* [f16; 2] (IEEE half)   <-> MIXED_PRECISION_F16_F32 */
} anufft_storage;

typedef enum anufft_profile {
  ANUFFT_PROFILE_HIGH_ACCURACY_F64 = 0,
  ANUFFT_PROFILE_LOW_PRECISION_F32 = 1,
  ANUFFT_PROFILE_MIXED_PRECISION_F16_F32 = 2
} anufft_profile;

anufft_status anufft_execute_type1_host(const anufft_plan *plan, int32_t storage,
                                        int32_t profile, int64_t num_points,
                                        const double *const coords[3], int64_t coord_stride,
                                        const void *values, void *out);
anufft_status anufft_execute_type2_host(const anufft_plan *plan, int32_t storage,
                                        int32_t profile, int64_t num_points,
                                        const double *const coords[3], int64_t coord_stride,
                                        const void *modes, void *out);

/* Exact direct sums on the device (replaces the WGPU direct kernels
 * nufft_type{1,2}_{1d,3d}, apollo-nufft-wgpu infrastructure/shaders/
 * nufft.wgsl:51-140, and their drivers NufftWgpuKernel::execute_type1_1d,
 * execute_type2_1d, execute_type1_3d, execute_type2_3d, kernel.rs:652-812,
 * reached through device.rs:112-216).  Over the plan's domain (n, L = n d):
 *   type1: out[k] = sum_j c_j exp(-2 pi i sum_a k_a x_{j,a} / L_a), all modes
 *          in FFT order ([nx][ny][nz], k_a signed: index <= n/2 -> index,
 *          else index - n);
 *   type2: out[j] = sum_k f_k exp(+2 pi i sum_a k_a x_{j,a} / L_a).
 * Device pointers: float32 coordinates (stride as for execute), complex64
 * values / modes; out is complex128 (interleaved doubles).  The reference
 * computes these sums in f32 and returns Complex64; here the arithmetic is
 * FP64.  No workspace; O(num_points * modes) work, queued on `stream`. */
anufft_status anufft_direct_type1(const anufft_plan *plan, int64_t num_points,
                                  const float *const coords[3], int64_t coord_stride,
                                  const float *values, double *out, void *stream);
anufft_status anufft_direct_type2(const anufft_plan *plan, int64_t num_points,
                                  const float *const coords[3], int64_t coord_stride,
                                  const float *modes, double *out, void *stream);

/* Exact Type-1 sum at chosen modes only: out[s] = sum_j c_j exp(-2 pi i
 * k(mode_index[s]).x_j / L), mode_index[s] a flat FFT-order index into
 * [nx][ny][nz] (1D: [n]).  Device pointers: float32 coordinates, complex128
 * values, int64 indices, complex128 out; FP64 arithmetic.  An extension of
 * the direct backend above (which always writes every mode) for sampled
 * accuracy checks of large problems; no workspace, queued on `stream`. */
anufft_status anufft_direct_type1_select(const anufft_plan *plan, int64_t num_points,
                                         const float *const coords[3], int64_t coord_stride,
                                         const double *values, int64_t num_modes,
                                         const int64_t *mode_index, double *out, void *stream);

/* The same exact sums from host pointers: FP64 coordinates (stride as for
 * execute) and complex128 values / modes, copied to the device unchanged;
 * FP64 arithmetic; complex128 out (host).  Replaces the CPU direct sums
 * nufft_type1_1d, nufft_type2_1d (apollo-nufft .../transform/
 * dimension_1d.rs:429-472) and nufft_type1_3d, nufft_type2_3d
 * (dimension_3d.rs:679-740) behind pyapollofft.nufft_type1_1d,
 * nufft_type2_1d and nufft_type1_3d (apollo-python/src/lib.rs:754-836).
 * Allocates device memory per call and synchronizes. */
anufft_status anufft_direct_type1_host(const anufft_plan *plan, int64_t num_points,
                                       const double *const coords[3], int64_t coord_stride,
                                       const double *values, double *out);
anufft_status anufft_direct_type2_host(const anufft_plan *plan, int64_t num_points,
                                       const double *const coords[3], int64_t coord_stride,
                                       const double *modes, double *out);

/* Per-stage timing with HIP events recorded on the execution stream
 * (no reference counterpart; the reference records no timings, every WGPU
 * pass sets timestamp_writes: None, kernel.rs:876).  While enabled, each
 * execute call records an event pair around each stage; collect waits for
 * them, returns the summed milliseconds and call counts per stage and clears
 * the records.  Not for concurrent use of one plan from several threads. */
typedef enum anufft_stage {
  ANUFFT_STAGE_FILL = 0,   /* zeroing of the oversampled grid */
  ANUFFT_STAGE_BIN = 1,    /* point binning (counting sort) and subproblems; with the
                              MFMA kernels also the per-point records */
  ANUFFT_STAGE_SPREAD = 2, /* Type-1 spreading */
  ANUFFT_STAGE_FFT = 3,    /* all FFT passes with pad/crop/deconvolution */
  ANUFFT_STAGE_INTERP = 4, /* Type-2 interpolation */
  ANUFFT_STAGE_GRAD = 5,   /* coordinate-gradient interpolation */
  ANUFFT_STAGE_PREP = 6,   /* per-point weight records (vector-ALU kernels) */
  ANUFFT_STAGE_TILE_MAX = 7, /* tile maxima of the grid (MFMA interpolation scales) */
  ANUFFT_STAGE_COUNT = 8
} anufft_stage;
anufft_status anufft_timing_enable(const anufft_plan *plan, int32_t on);
anufft_status anufft_timing_collect(const anufft_plan *plan, double *ms, int64_t *calls);

/* Diagnostics for stage-level parity tests (no reference counterpart).
 * anufft_debug_sort: the point binning of a 3D plan -- perm[i] is the input
 * index of the i-th point in tile order (stable: ties keep input order),
 * bin_start[b] (nbins + 1 entries) the first sorted position of tile b.
 * anufft_debug_spread: the oversampled grid after Type-1 spreading
 * (mx*my*mz complex, before the FFT), the quantity apollo-nufft accumulates
 * in type1_into (dimension_3d.rs:277-327). */
anufft_status anufft_debug_num_bins(const anufft_plan *plan, int64_t *num_bins);
anufft_status anufft_debug_sort(const anufft_plan *plan, int64_t num_points,
                                const float *const coords[3], int64_t coord_stride,
                                uint32_t *perm, uint32_t *bin_start, void *workspace,
                                size_t workspace_bytes, void *hip_stream);
anufft_status anufft_debug_spread(const anufft_plan *plan, int64_t num_points,
                                  const float *const coords[3], int64_t coord_stride,
                                  const float *values, float *grid, void *workspace,
                                  size_t workspace_bytes, void *hip_stream);

/* Message of the last failed call on this thread ("" if none). */
const char *anufft_last_error(void);

/* Library version string. */
const char *anufft_version(void);

#ifdef __cplusplus
}
#endif

#endif /* APOLLO_NUFFT_H */
