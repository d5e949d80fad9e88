/*
 * lem.h -- C-ABI of liblem.so, the gfx950 (MI355X) library behind the LyceanEM
 * raycast + frequency-domain propagation path.
 *
 * The reference binds this path two ways: numba.cuda kernels launched from
 * Python (kernel1Dv3, freqdomainkernal) and the pybind11 module `lyceanem.em`
 * (lyceanem/src/frequency_domain_models.cu:314-320).  This header replaces both
 * with plain `extern "C"` entry points that ctypes (or any FFI) can bind.
 *
 * Conventions
 *   - Every buffer is a caller-owned DEVICE pointer (e.g. torch.Tensor.data_ptr()
 *     of a contiguous tensor on the current HIP device).  Points are N x 3 float32
 *     (x, y, z rows, like the reference's `unified_model`), complex weights are
 *     N x 3 complex64 (interleaved re, im), triangles are T x 9 float32
 *     (v0, v1, v2 -- the layout of base_types.triangle_t, base_types.py:21-38).
 *   - `stream` is a hipStream_t (NULL = the default stream); all work is enqueued
 *     on it and the call returns without synchronising.
 *   - Return value: 0 on success, a negative LEM_E_* code otherwise, with a
 *     message in lem_last_error() (thread-local).  Nothing here calls exit() or
 *     abort(), unlike gpuErrchk (lyceanem/src/gpu_error.cuh:2-10).
 *   - No entry point allocates device memory; scratch space comes from a
 *     caller-supplied workspace whose size a *_workspace() query returns.
 */
#ifndef LEM_H
#define LEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LEM_ABI_VERSION 1

#define LEM_OK 0
#define LEM_E_INVALID (-1)   /* bad argument (sizes, null pointers, flags) */
#define LEM_E_HIP (-2)       /* HIP runtime or launch error */
#define LEM_E_WORKSPACE (-3) /* workspace missing or too small */

/* Predicate semantics.
 * LEM_SEM_NUMBA : numba `hit`, rayfunctions.py:37-104 -- two-sided,
 *                 EPSILON = 1e-7, float32 edges/cross/dot, float64 F, U, V, t,
 *                 accept window (2e-7, dist - 2e-7), kernel1Dv3's shrinking
 *                 closest-hit `dist` (rayfunctions.py:212-244).
 * LEM_SEM_NATIVE: native `hit`, lyceanem/src/raycasting.cuh:19-68 -- back-face
 *                 culled, EPSILON = 1e-6f, all float32, any-hit only. */
#define LEM_SEM_NUMBA 0
#define LEM_SEM_NATIVE 1

/* Visibility flags. */
#define LEM_VIS_ANY_HIT 1   /* stop each ray at its first accepted hit (mask only) */
#define LEM_VIS_SKIP_SELF 2 /* never test pairs with source row == sink row
                               (native flag = !self_to_self,
                               raycasting_accelerated.cuh:72) */
#define LEM_VIS_SINK_MAJOR 4 /* grid traversal only: lanes run over sources, bits
                                are K x ceil(S/64) words and dist is K x S */
#define LEM_VIS_INPUT_ORDER 8 /* grid traversal only: wavefronts take the lane-side
                                 points (sources when sink-major, else sinks) in
                                 input order instead of a spatial (Morton) order;
                                 for inputs already ordered in compact groups of 64 */
#define LEM_VIS_FROM_SINK 16 /* grid traversal, any-hit only: walk each ray's cells from
                                its sink end (the same segment; every triangle is
                                still tested with the source -> sink ray, so the
                                mask is unchanged); finds blockers that sit near the
                                sinks -- e.g. a scatterer's own surface -- first */

/* Field flags. */
#define LEM_FIELD_SINK_TANGENT 1 /* project onto the sink tangent plane, as
                                    freqdomainkernal does for line-of-sight rows
                                    of a 3-column index (empropagation.py:1150-1161) */
#define LEM_FIELD_ELEMENTS 2     /* write the S x K x 3 network instead of the
                                    source sum (frequency_domain.py:703-716) */
#define LEM_FIELD_BITS_SINK_MAJOR 4 /* hit_bits is K x ceil(S/64) words (bit s%64 of
                                       word s/64 of row k), the layout
                                       lem_visibility_grid writes with
                                       LEM_VIS_SINK_MAJOR */
#define LEM_FIELD_INPUT_ORDER 8     /* lem_field_los_grid (sinks) and lem_field_los_adjoint
                                       (sources): lanes take the points in input order
                                       instead of a spatial (Morton) order */

int lem_abi_version(void);
const char *lem_last_error(void);
/* Writes the current device's gcnArchName (e.g. "gfx950:sramecc+:xnack-"). */
int lem_device_arch(char *buf, int len);

/* Dense S x K brute-force visibility.
 * Replaces the first raycast wave of workchunkingv2 (rayfunctions.py:2681-2783:
 * create_model_index -> charge_rays_environment1Dv2 -> kernel1Dv3 -> rayHits1Dv2)
 * for source -> sink rays, without materialising the S*K index or ray payload.
 * src: S x 3, snk: K x 3, tris: T x 9.
 * hit_bits: S x ceil(K/64) uint64 words; bit k%64 of word k/64 of row s is the
 *   kernel1Dv3 `intersect` flag of ray s -> k (1 = blocked).
 * dist: S x K float32 kernel1Dv3 `dist` (NULL to skip; needs the closest-hit
 *   mode with LEM_SEM_NUMBA, i.e. flags without LEM_VIS_ANY_HIT). */
int lem_visibility_dense(const float *src, int64_t S, const float *snk, int64_t K,
                         const float *tris, int64_t T, int semantics, int flags,
                         uint64_t *hit_bits, float *dist, void *stream);

/* Explicit ray list: the kernel1Dv3 launch of launchRaycaster1Dv3 /
 * chunkingRaycaster1Dv3 (rayfunctions.py:1951-2017, 2102-2369) including the
 * ray charging of charge_rays_environment1Dv2 (rayfunctions.py:1790-1842).
 * points: N x 3 unified point table; pairs: R x 2 int32 0-based
 * (origin row, target row).  intersect: R bytes (0/1, 0xFF for a pair outside
 * the table); dist: R float32 or NULL. */
int lem_raycast_pairs(const float *points, int64_t n_points, const int32_t *pairs,
                      int64_t n_rays, const float *tris, int64_t n_tris, int semantics,
                      int flags, uint8_t *intersect, float *dist, void *stream);

/* Line-of-sight field, fused over the visibility mask.
 * Replaces freqdomainkernal (empropagation.py:1083-1265) for 2-column paths
 * (and the line-of-sight rows of a 3-column index) plus the host source sum
 * `np.dot(ones(S), scatter_map[:, :, c])` (frequency_domain.py:647-649).
 * src: S x 3, src_w: S x 3 complex64; snk, snk_n: K x 3; snk_w: K x 3 complex64
 * or NULL (weight 1, frequency_domain.py:398-400); hit_bits as produced by
 * lem_visibility_dense (S x ceil(K/64)), or K x ceil(S/64) with
 * LEM_FIELD_BITS_SINK_MAJOR, or NULL (all visible).
 * out: K x 3 complex64 (source sum), or S x K x 3 complex64 with
 * LEM_FIELD_ELEMENTS.  workspace: lem_field_los_workspace() bytes. */
int64_t lem_field_los_workspace(int64_t S, int64_t K, int flags);
int lem_field_los(const float *src, const float *src_w, int64_t S, const float *snk,
                  const float *snk_n, const float *snk_w, int64_t K, const uint64_t *hit_bits,
                  double alpha, double beta, int flags, float *out, void *workspace,
                  int64_t workspace_bytes, void *stream);

/* lem_field_los without a workspace and without any allocation, for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  The source-sum partials live in a static device array of
 * LEM_FIELD_LOS_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends);
 * K above lem_field_los_noalloc_max_sinks() is rejected with LEM_E_INVALID.
 * Every call on one GPU shares that array: calls that overlap on one GPU (two
 * streams, two threads) race and give wrong sums.  The result equals
 * lem_field_los's; when 12 * K * (its chunk count) floats exceed the array,
 * fewer source chunks are used and the float32 partial sums are grouped
 * differently (same float64 merge).  Arguments as lem_field_los, minus the
 * workspace; LEM_FIELD_ELEMENTS needs no scratch and is passed through. */
#define LEM_FIELD_LOS_NOALLOC_SCRATCH_BYTES (16ll << 20)
int64_t lem_field_los_noalloc_max_sinks(void);
int lem_field_los_noalloc(const float *src, const float *src_w, int64_t S, const float *snk,
                          const float *snk_n, const float *snk_w, int64_t K,
                          const uint64_t *hit_bits, double alpha, double beta, int flags,
                          float *out, void *stream);

/* Generic path list: the network EMGPUFreqDomain returns
 * (empropagation.py:2197-2410).  pos, nrm: N x 3; w: N x 3 complex64 (the
 * scattering_t table, base_types.py:89-119); index: V x ncol int32, 1-based,
 * 0 = absent point (workchunkingv2's full_index).  Paths are grouped by their
 * (source, sink) slot from targettingindex (empropagation.py:3394-3429).
 * eta (N float32, or NULL when every point is electric): eta[i] > 0 marks a
 * magnetic point (Electric = False) of impedance sqrt(Re permeability /
 * Re permittivity); a path starting at one launches (w x d) / eta, d the unit
 * direction of its first leg, instead of w (empropagation.py:1111-1146):
 * `order` lists the path rows slot by slot, seg_start (nseg + 1) bounds each
 * slot's run, seg_slot = s * K + k.  out: S x K x 3 complex128, only the listed
 * slots are written (zero it first).  err_flag (device int, zeroed by the caller)
 * is set to 1 if a row points outside the table.  ncol 2 to 4: a zero entry at
 * column i >= 1 reads the last point of the table, as freqdomainkernal's
 * negative index does under numba (a line-of-sight row [s, k, 0, 0] of the
 * 4-column index is projected onto that point's tangent plane after the
 * sink's). */
int lem_field_paths(const float *pos, const float *nrm, const float *w, const float *eta,
                    int64_t N, const int32_t *index, int ncol, const int64_t *order,
                    const int64_t *seg_start, const int64_t *seg_slot, int64_t nseg,
                    double alpha, double beta, double *out, int *err_flag, void *stream);

/* Uniform-grid acceleration structure (host-side descriptor; the arrays it
 * describes are device buffers the caller allocates from the sizes here). */
typedef struct {
    float lo[3];    /* grid origin: scene bounds minus 2*eps */
    float cell[3];  /* cell edge per axis */
    int32_t n[3];   /* cells per axis */
    int32_t mshift; /* macro blocks of 2^mshift cells per axis (empty-space skipping) */
    int64_t ncells; /* n[0]*n[1]*n[2]; cell (x,y,z) is ((z*n[1] + y)*n[0] + x) */
    int64_t nrefs;  /* triangle references summed over cells */
    float eps;      /* inflation of triangle boxes and traversal margin */
    int32_t nmacro; /* macro blocks; their boxes follow the ncells cell boxes */
} lem_grid_info;

/* Grid build, replacing the tile binning of bin_counts_to_numpy /
 * bin_triangles_to_numpy (terrain_acceleration_build.cuh:287-431) with a
 * general 3-D grid.  cells_per_tri sets the cell budget (cells ~ T * value,
 * 2 is a good default).
 * 1) lem_grid_plan: bounds, resolution and per-cell counts; synchronises the
 *    stream and fills *info (ncells, nrefs).
 * 2) the caller allocates cell_start (ncells + 1 int32), tri_ids (nrefs int32),
 *    recs (nrefs x 12 float32: v0, e1, e2, the triangle id as int32 bits,
 *    padding; cell by cell, ids ascending within a cell) and
 *    cell_box ((ncells + nmacro) x 8 + nmacro x 2 float32: the tight box of
 *    each cell's contents, lo.xyz, start, hi.xyz, end -- start/end being
 *    cell_start[c] and cell_start[c + 1] as int32 bits --, then each macro
 *    block's box (w lanes 0), then each macro block's box quantised to 8
 *    bytes), and
 * 3) lem_grid_fill writes them; the same workspace must be passed to both. */
int64_t lem_grid_workspace(int64_t T, float cells_per_tri);
int lem_grid_plan(const float *tris, int64_t T, float cells_per_tri, lem_grid_info *info,
                  void *workspace, int64_t workspace_bytes, void *stream);
int lem_grid_fill(const float *tris, int64_t T, const lem_grid_info *info, int32_t *cell_start,
                  int32_t *tri_ids, float *recs, float *cell_box, void *workspace,
                  int64_t workspace_bytes, float cells_per_tri, void *stream);

/* Dense S x K visibility through the grid (DDA): the same masks (and, with
 * LEM_SEM_NUMBA, kernel1Dv3 distances) as lem_visibility_dense with the same
 * semantics.  kernel1Dv3's distance depends on the order of near-tied hits
 * (its window shrinks to 2e-7 below each accepted hit); the grid replays that
 * window in triangle-id order over the hits near the closest one, so distances
 * match the brute-force order too (exact unless more than 8 hits chain within
 * 2e-7 + one float32 ulp of each other).  Replaces the tile traversal raycast_tiles / DDA
 * (raycasting_accelerated.cuh:106-268, 458-482) for source -> sink rays.
 * Flags: LEM_VIS_ANY_HIT, LEM_VIS_SINK_MAJOR, LEM_VIS_SKIP_SELF.
 * LEM_SEM_NATIVE is any-hit (no distance).  Unless LEM_VIS_INPUT_ORDER is
 * set, the lane-side points are first sorted into a 3-D Morton order so that
 * each wavefront's 64 rays are spatially compact; results are written at the
 * original indices.  With the persistent traversal (macro table in LDS) and
 * rows of at most 524,288 lane-side points, each wavefront stores its 64 bits
 * as one word in lane order and the rows are reordered in place afterwards;
 * otherwise hit_bits is cleared and bits are set atomically.  The persistent
 * traversal also claims rows in a Morton order of the other side's points.
 * The workspace (lem_visibility_grid_workspace(S, K, flags) bytes) holds the
 * task counters of the persistent traversal (one per XCD: each XCD walks a
 * contiguous eighth of every row of rays, so its L2 keeps one region's cells)
 * and both orders and their sorts: calls that overlap on one workspace race. */
int64_t lem_visibility_grid_workspace(int64_t S, int64_t K, int flags);
int lem_visibility_grid(const float *src, int64_t S, const float *snk, int64_t K,
                        const lem_grid_info *info, const int32_t *cell_start, const float *recs,
                        const float *cell_box, int semantics, int flags, uint64_t *hit_bits,
                        float *dist, void *workspace, int64_t workspace_bytes, void *stream);

/* Line-of-sight field fused with the grid traversal: lem_field_los (source
 * sum only) where each ray is cast any-hit through the grid instead of read
 * from a precomputed mask -- the S x K mask is never stored.  Lanes run over
 * the sinks in a Morton order unless LEM_FIELD_INPUT_ORDER is set (each sink's
 * sum, and so the result, is the same).  workspace:
 * lem_field_los_grid_workspace(S, K, flags) bytes. */
int64_t lem_field_los_grid_workspace(int64_t S, int64_t K, int flags);
int lem_field_los_grid(const float *src, const float *src_w, int64_t S, const float *snk,
                       const float *snk_n, const float *snk_w, int64_t K,
                       const lem_grid_info *info, const int32_t *cell_start, const float *recs,
                       const float *cell_box, double alpha, double beta, int flags, float *out,
                       void *workspace, int64_t workspace_bytes, void *stream);

/* Hand-written adjoint of lem_field_los (source sum) with the mask frozen.
 * grad_out: K x 3 complex64, dLoss/dE in PyTorch's convention
 * (dL/dRe E + j dL/dIm E).  Writes grad_w (S x 3 complex64, dLoss/dw_s) and
 * grad_p (S x 3 float32, dLoss/d source position); either may be NULL.
 * flags: the forward's LEM_FIELD_SINK_TANGENT; snk_w as in the forward.  Lanes
 * run over the sources in a Morton order unless LEM_FIELD_INPUT_ORDER is set
 * (the per-source sums, and so the result, are the same).
 * No counterpart in the reference (it has no gradients); it differentiates
 * lossy_propagation and the launch projector (empropagation.py:178-249,
 * 768-801) as freqdomainkernal combines them for 2-column paths. */
int64_t lem_field_los_adjoint_workspace(int64_t S, int64_t K);
int lem_field_los_adjoint(const float *src, const float *src_w, int64_t S, const float *snk,
                          const float *snk_n, const float *snk_w, int64_t K,
                          const uint64_t *hit_bits, double alpha, double beta, int flags,
                          const float *grad_out, float *grad_w, float *grad_p, void *workspace,
                          int64_t workspace_bytes, void *stream);

/* dB mask loss of a source-summed far field and its gradient, in one pass:
 * P_k = sum_c |E_kc|^2, R = ref_power if > 0 else max_k P_k (a constant of the
 * loss), db_k = 10 log10(P_k / R + eps), err_k = db_k for in_beam[k] != 0 and
 * max(db_k - sidelobe_db, 0) otherwise; *loss = mean_k err_k^2 (one float on
 * the device).  grad_E (K x 3 complex64, or NULL) receives dloss/dE in
 * PyTorch's convention (dL/dRe E + j dL/dIm E).  E: K x 3 complex64; in_beam:
 * K bytes.  workspace: lem_db_mask_loss_workspace(K) bytes.  Deterministic,
 * no host synchronisation and no memset (capturable in a HIP graph).
 * No counterpart in the reference (it has no gradients): the beam-synthesis
 * objective for lem_field_los / lem_field_los_adjoint. */
int64_t lem_db_mask_loss_workspace(int64_t K);
int lem_db_mask_loss(const float *E, const uint8_t *in_beam, int64_t K, double sidelobe_db,
                     double ref_power, double eps, float *loss, float *grad_E, void *workspace,
                     int64_t workspace_bytes, void *stream);

/* Batched-excitation contraction on FP32 MFMA: for F wavenumbers k_f and B = 64
 * excitation sets W at once,
 *   E[f,c,k,b] = sum_s G_c[k,s,f] W[f,s,b],
 *   G[k,s,f]   = F_f(r) (d . n_k) (e - d (d . e)),  F_f the lossy_propagation
 *                factor at wavenumber k_f (empropagation.py:768-801), e = (pol_x,
 *                pol_y, pol_z) the source polarisation, every path visible.
 * Equals B x F calls of lem_field_los with weights W[f,:,b] e (the reference
 * runs one calculate_scattering per excitation; its multiE branch cannot run,
 * frequency_domain.py:612-617).  src: S x 3 (S a multiple of 32), snk, snk_n:
 * K x 3 (K a multiple of 32), wavenumbers: F float64, W: F x S x 64 complex64,
 * out: F x 3 x K x 64 complex64. */
int lem_gw_mfma(const float *src, int64_t S, float pol_x, float pol_y, float pol_z,
                const float *snk, const float *snk_n, int64_t K, const double *wavenumbers,
                int64_t F, double alpha, const float *W, int64_t B, float *out, void *stream);

/* Instrumented grid traversal for measurement: adds, over all S x K rays, the
 * Moller-Trumbore tests performed (stats[0]), cells of live macro blocks visited
 * (stats[1]), blocked rays (stats[2]), macro blocks entered (stats[3]) and the
 * tests a conservative front-to-back DDA would perform up to the first hit (the
 * triangle references of every crossed cell, no pruning; stats[4]) to the
 * device uint64 array `stats` of 5 entries (zero it first).
 * Writes no masks.  No counterpart in the reference (it reports no work counts). */
int lem_visibility_grid_stats(const float *src, int64_t S, const float *snk, int64_t K,
                              const lem_grid_info *info, const int32_t *cell_start,
                              const float *recs, const float *cell_box, int flags,
                              uint64_t *stats, void *stream);

/* Projected-area pattern of visiblespace (rayfunctions.py:316-427; the
 * angle_pop / patternsort loops, :431-468): out[k] (float32) = sum over the
 * sources s whose ray to shell point k is unblocked (hit_bits: source-major
 * S x ceil(K/64) from lem_visibility_dense) of area[s] * cos(angle between the
 * ray and the normal of s), angles over 90 degrees contributing nothing;
 * area == NULL gives every visible source a portion of 1 (the reference's
 * vertex_area == 0).  Sources are summed in index order in float64. */
int lem_projected_area(const float *src, const float *nrm, const float *area, int64_t S,
                       const float *snk, int64_t K, const uint64_t *hit_bits, float *out,
                       void *stream);

/* Beamformed-pattern power of many weight rows on FP32 MFMA, the hot loop of
 * MaximumDirectivityMap (beamforming.py:484-673): for each row b of W
 * (nb x ne complex64), B = W[b] . Et and W[b] . Ep over the element patterns
 * (ne x np complex64, points el-major / az-fastest); power[b] (float64) =
 * sum_p q[p] (|Bt|^2 + |Bp|^2) and ustar[b] = (|Bt|^2, |Bp|^2) at point
 * pstar[b] (float32; pstar < 0: not written).  nb a multiple of 32, ne of 32
 * (zero-pad), np of 64 (zero-pad q).  Workspace: lem_beam_workspace bytes.
 * No counterpart kernel in the reference (its GPUBeamformingMap is a stub,
 * beamforming.py:1553-1562; the work is a numpy loop there). */
int64_t lem_beam_workspace(int64_t nb, int64_t np);
int lem_beam_power(const float *W, int64_t nb, const float *Et, const float *Ep, int64_t ne,
                   int64_t np, const float *q, const int32_t *pstar, double *power, float *ustar,
                   void *workspace, int64_t workspace_bytes, void *stream);

/* ---- time domain (models/time_domain.py, TimeDomainv3) ------------------- */

/* Per path row of a 1-based index (V x ncol int32, 0 = unused column, the
 * reference's negative-index wrap reproduced): the real polarisation vector
 * carried along the path scaled by |product of the lossy_propagation factors|
 * (amp, V x 3 float64) and the arrival time path length / c + the imaginary
 * parts of the point weights (arrival, V float64).  Points: N x 3 float32 with
 * normals, weights N x 3 complex64.  Replaces the per-path part of
 * timedomainkernal (empropagation.py:1345-1497).  eta as for
 * lem_field_paths: a magnetic first point launches Re(w) x d / eta
 * (:1395-1419).
 * A row whose first entry is 0 or with an entry outside [0, N] sets *err
 * (device int32, zero it first) and gets NaN outputs. */
int lem_time_paths(const float *pos, const float *nrm, const float *w, const float *eta,
                   int64_t N, const int32_t *index, int64_t V, int ncol, double alpha, double beta,
                   double *amp, double *arrival, int32_t *err, void *stream);

/* *wake (device float64) = min(1 s, min arrival): the wake time of
 * timedomainkernal's atomic minimum (empropagation.py:1490-1495), taken over
 * every path (the reference's read after a block barrier is launch-order
 * dependent). */
int lem_time_wake(const double *arrival, int64_t V, double *wake, void *stream);

/* Start sample of each path: (arrival - *wake) // (1 / sampling_freq) with
 * Python's float floor division (empropagation.py:1503-1506), as int64. */
int lem_time_bins(const double *arrival, int64_t V, const double *wake, double sampling_freq,
                  int64_t *bin, void *stream);

/* Time map per (source, sink) slot: out[slot][n][c] (float64, N samples) =
 * sum over the slot's paths of excitation[n - bin] * amp[c] for
 * 0 <= n - bin < E -- timedomainkernal's atomic placement
 * (empropagation.py:1507-1600) as a gather in path order.  Paths grouped by
 * slot as for lem_field_paths (order, seg_start[nseg + 1], seg_slot[nseg]);
 * slots without paths are not written. */
int lem_time_map(const double *amp, const int64_t *bin, const int64_t *order,
                 const int64_t *seg_start, const int64_t *seg_slot, int64_t nseg,
                 const double *excitation, int64_t E, int64_t N, double *out, void *stream);

/* The time map summed over every path: out[n][c] (N x 3 float64), the source
 * and sink sums of time_domain.calculate_scattering (time_domain.py:383-428).
 * Paths are sorted stably by start sample, their amplitudes summed per sample
 * in path order and convolved with the excitation.  V and N below 2^31; the
 * workspace size comes from lem_time_trace_workspace (needs a HIP device: the
 * sort's scratch size depends on it). */
int64_t lem_time_trace_workspace(int64_t V, int64_t N);
int lem_time_trace(const double *amp, const int64_t *bin, int64_t V, const double *excitation,
                   int64_t E, int64_t N, double *out, void *workspace, int64_t workspace_bytes,
                   void *stream);

#ifdef __cplusplus
}
#endif

#endif /* LEM_H */
