// lem_fields.hip -- fused, visibility-masked polarised Rayleigh-Sommerfeld
// propagation for gfx950.
//
// Replaces freqdomainkernal (empropagation.py:1083-1265) behind
// EMGPUFreqDomain (empropagation.py:2197-2410) and the native em_wave /
// ray_launch pair (em.cuh:18-120).  Per (source s, sink k) the contribution is
//
//     L_sk * (I - d d^T) w_s          d = (p_k - p_s)/r
//     L_sk = -(1/2pi) (-(alpha + j beta) - 1/r) e^{-(alpha + j beta) r}/r (d . n_k)
//
// then multiplied componentwise by the sink weight and (for the 3-column index
// of calculate_scattering with scattering > 0) projected onto the sink tangent
// plane -- both linear and source-independent, so the source sum applies them
// once per sink.  r and the phase are float64; the rest is float32, summed over
// groups of 16 sources whose partials are Kahan-accumulated; chunk partials
// merge in a fixed order, so results are deterministic (the reference adds
// with float64 atomics in arbitrary order).
#include "lem_device.h"
#include "lem_field_dev.h"
#include "lem_internal.h"
#include "../../include/lem.h"

#include <mutex>

namespace lem {

// Scratch of lem_field_los_noalloc: a file-scope device array, one copy per GPU
// the library runs on, held until the process ends.  Host code never passes
// its name to a kernel (on the host it is only a stand-in for the symbol): the
// device address comes from hipGetSymbolAddress, once per device.
constexpr int64_t NOALLOC_SCRATCH_FLOATS = LEM_FIELD_LOS_NOALLOC_SCRATCH_BYTES / 4;
static __device__ __align__(256) float scratch[NOALLOC_SCRATCH_FLOATS];
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static float *g_scratch_ptr[MAX_DEVICES];

// Source-summed partials: block b covers sinks [(b % nbk)*256, +256) and
// source chunk c = b / nbk ([c*CH, min(S, (c+1)*CH))).
template <bool LOSSY>
__global__ __launch_bounds__(FB) void k_field_los_partial(
    const float *__restrict__ src, const float *__restrict__ src_w, int64_t S,
    const float *__restrict__ snk, const float *__restrict__ snk_n, int64_t K,
    const unsigned long long *__restrict__ hit_bits, int64_t W64, int sink_major, float alpha,
    double beta, int64_t CH, int64_t nbk, float *__restrict__ ws_sum, float *__restrict__ ws_cmp)
{
    __shared__ SrcLds s_src[FS_TILE];
    // source-major mask: the tile's words for this block's 256 sinks (4 per
    // source), read from LDS in the loop instead of one global load per source
    __shared__ unsigned long long s_mask[FS_TILE][FB / 64];
    const int64_t c = (int64_t)blockIdx.x / nbk;
    const int64_t k0 = ((int64_t)blockIdx.x % nbk) * FB;
    const int64_t k = k0 + threadIdx.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const bool active = k < K;
    const int64_t kk = active ? k : 0;
    const double xk = snk[3 * kk], yk = snk[3 * kk + 1], zk = snk[3 * kk + 2];
    const float nkx = snk_n[3 * kk], nky = snk_n[3 * kk + 1], nkz = snk_n[3 * kk + 2];
    const int64_t s0 = c * CH;
    const int64_t s1 = (s0 + CH) < S ? (s0 + CH) : S;
    const bool src_major_bits = hit_bits && !sink_major;
    KahanC3 acc;
    acc.zero();
    // sink-major mask: one word covers 64 consecutive sources of this lane's
    // row; every lane walks the same sources, so the reload is wave-uniform
    unsigned long long word = 0;
    int64_t word_at = -1;
    for (int64_t base = s0; base < s1; base += FS_TILE) {
        __syncthreads();
        stage_sources(s_src, src, src_w, base, s1);
        if (src_major_bits) {
            const int64_t sj = base + threadIdx.x;
#pragma unroll
            for (int i = 0; i < FB / 64; ++i) {
                const int64_t wi = (k0 >> 6) + i;
                s_mask[threadIdx.x][i] = (sj < s1 && wi < W64) ? hit_bits[sj * W64 + wi] : 0ull;
            }
        }
        __syncthreads();
        int n = (int)((s1 - base) < FS_TILE ? (s1 - base) : FS_TILE);
        if (active) {
            for (int j0 = 0; j0 < n; j0 += SUM_GROUP) {
                const int j1 = (j0 + SUM_GROUP) < n ? (j0 + SUM_GROUP) : n;
                float g[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 2
                for (int j = j0; j < j1; ++j) {
                    const int64_t s = base + j;
                    bool bl = false;
                    if (sink_major) {
                        if ((s >> 6) != word_at) {
                            word_at = s >> 6;
                            word = hit_bits ? hit_bits[k * W64 + word_at] : 0ull;
                        }
                        bl = (word >> (s & 63)) & 1ull;
                    } else if (src_major_bits) {
                        bl = (s_mask[j][wave] >> lane) & 1ull;
                    }
                    if (bl)
                        continue;
                    const SrcLds &L = s_src[j];
                    C3 w;
                    w.xr = L.w[0]; w.xi = L.w[1]; w.yr = L.w[2]; w.yi = L.w[3]; w.zr = L.w[4];
                    w.zi = L.w[5];
                    C3 e;
                    float lr, li;
                    los_leg<LOSSY>(L.x, L.y, L.z, w, xk, yk, zk, nkx, nky, nkz, alpha, beta, e,
                                   lr, li);
                    add_leg(g, e, lr, li);
                }
                acc.add(g);
            }
        }
    }
    if (active) {
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            ws_sum[(c * 6 + i) * K + k] = acc.s[i];
            ws_cmp[(c * 6 + i) * K + k] = acc.c[i];
        }
    }
}

// Fixed-order merge of the chunk partials, then sink weight and tangent step.
// out: K x 3 complex64 (interleaved).
__global__ __launch_bounds__(FB) void k_field_los_finish(
    const float *__restrict__ ws_sum, const float *__restrict__ ws_cmp, int64_t NC, int64_t K,
    const float *__restrict__ snk_n, const float *__restrict__ snk_w, int sink_tangent,
    float *__restrict__ out)
{
    const int64_t k = (int64_t)blockIdx.x * FB + threadIdx.x;
    if (k >= K)
        return;
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t c = 0; c < NC; ++c) {
#pragma unroll
        for (int i = 0; i < 6; ++i)
            a[i] += (double)ws_sum[(c * 6 + i) * K + k] - (double)ws_cmp[(c * 6 + i) * K + k];
    }
    C3 e;
    e.xr = (float)a[0]; e.xi = (float)a[1]; e.yr = (float)a[2];
    e.yi = (float)a[3]; e.zr = (float)a[4]; e.zi = (float)a[5];
    if (snk_w)
        e = cmul_comp(e, load_w(snk_w, k));
    if (sink_tangent)
        e = project_normal(e, snk_n[3 * k], snk_n[3 * k + 1], snk_n[3 * k + 2]);
    float2 *o = reinterpret_cast<float2 *>(out) + 3 * k;
    o[0] = make_float2(e.xr, e.xi);
    o[1] = make_float2(e.yr, e.yi);
    o[2] = make_float2(e.zr, e.zi);
}

// Per-element network (calculate_scattering(elements=True)): out is
// S x K x 3 complex64; block b covers source s = b / nbk, 256 sinks.
template <bool LOSSY>
__global__ __launch_bounds__(FB) void k_field_los_elements(
    const float *__restrict__ src, const float *__restrict__ src_w, const float *__restrict__ snk,
    const float *__restrict__ snk_n, const float *__restrict__ snk_w, int64_t K,
    const unsigned long long *__restrict__ hit_bits, int64_t W64, int sink_major, float alpha,
    double beta, int sink_tangent, int64_t nbk, float *__restrict__ out)
{
    const int64_t s = (int64_t)blockIdx.x / nbk;
    const int64_t k = ((int64_t)blockIdx.x % nbk) * FB + threadIdx.x;
    if (k >= K)
        return;
    C3 e;
    if (blocked_any(hit_bits, W64, sink_major != 0, s, k)) {
        e.xr = e.xi = e.yr = e.yi = e.zr = e.zi = 0.0f;
    } else {
        const float nkx = snk_n[3 * k], nky = snk_n[3 * k + 1], nkz = snk_n[3 * k + 2];
        e = los_term<LOSSY>(src[3 * s], src[3 * s + 1], src[3 * s + 2], load_w(src_w, s), snk[3 * k],
                     snk[3 * k + 1], snk[3 * k + 2], nkx, nky, nkz, alpha, beta);
        if (snk_w)
            e = cmul_comp(e, load_w(snk_w, k));
        if (sink_tangent)
            e = project_normal(e, nkx, nky, nkz);
    }
    float2 *o = reinterpret_cast<float2 *>(out) + 3 * (s * K + k);
    o[0] = make_float2(e.xr, e.xi);
    o[1] = make_float2(e.yr, e.yi);
    o[2] = make_float2(e.zr, e.zi);
}

// Generic path list (EMGPUFreqDomain's input): one thread per (source, sink)
// slot, summing that slot's paths in index order.  index rows are 1-based,
// 0 = absent point; `order` lists path rows grouped by slot.  eta (nullable):
// eta[i] > 0 marks a magnetic point, whose paths launch (w x d) / eta.
template <bool LOSSY>
__global__ __launch_bounds__(FB) void k_field_paths(
    const float *__restrict__ pos, const float *__restrict__ nrm, const float *__restrict__ w,
    const float *__restrict__ eta, int64_t N, const int32_t *__restrict__ index, int ncol,
    const int64_t *__restrict__ order,
    const int64_t *__restrict__ seg_start, const int64_t *__restrict__ seg_slot, int64_t nseg,
    float alpha, double beta, double *__restrict__ out, int *__restrict__ err)
{
    const int64_t g = (int64_t)blockIdx.x * FB + threadIdx.x;
    if (g >= nseg)
        return;
    KahanC3 acc;
    acc.zero();
    for (int64_t q = seg_start[g]; q < seg_start[g + 1]; ++q) {
        const int32_t *row = index + (int64_t)ncol * order[q];
        bool ok = row[0] >= 1 && row[0] <= N && row[1] >= 1 && row[1] <= N;
        for (int i = 2; i < ncol; ++i)
            ok = ok && row[i] >= 0 && row[i] <= N;
        if (!ok) {
            *err = 1;
            continue;
        }
        C3 E = load_w(w, row[0] - 1);
        float lr = 1.0f, li = 0.0f;
        for (int i = 0; i < ncol - 1; ++i) {
            // a zero entry (no point) reads the last point of the table, as
            // freqdomainkernal's network_index - 1 = -1 does under numba's
            // negative-index wrap (empropagation.py:1147-1156): line-of-sight
            // rows of a 4-column index are projected onto that point's tangent
            // plane after the sink's
            int64_t a = row[i] > 0 ? row[i] - 1 : N - 1;
            int32_t b = row[i + 1];
            if (i >= 1)
                E = project_normal(E, nrm[3 * a], nrm[3 * a + 1], nrm[3 * a + 2]);
            if (b != 0) {
                int64_t bb = b - 1;
                double dx = (double)pos[3 * bb] - (double)pos[3 * a];
                double dy = (double)pos[3 * bb + 1] - (double)pos[3 * a + 1];
                double dz = (double)pos[3 * bb + 2] - (double)pos[3 * a + 2];
                Phase p = phase_of(dx, dy, dz, beta);
                float ux = (float)dx * p.inv_r, uy = (float)dy * p.inv_r,
                      uz = (float)dz * p.inv_r;
                float cosang = ux * nrm[3 * bb] + uy * nrm[3 * bb + 1] + uz * nrm[3 * bb + 2];
                float l2r, l2i;
                leg_factor<LOSSY>(p, alpha, (float)beta, cosang, l2r, l2i);
                if (i == 0 && eta && eta[a] > 0.0f) {
                    // magnetic source: (w x d) / eta
                    const float ie = 1.0f / eta[a];
                    C3 m;
                    m.xr = (E.yr * uz - E.zr * uy) * ie;
                    m.xi = (E.yi * uz - E.zi * uy) * ie;
                    m.yr = (E.zr * ux - E.xr * uz) * ie;
                    m.yi = (E.zi * ux - E.xi * uz) * ie;
                    m.zr = (E.xr * uy - E.yr * ux) * ie;
                    m.zi = (E.xi * uy - E.yi * ux) * ie;
                    E = m;
                }
                float nr = lr * l2r - li * l2i;
                float ni = lr * l2i + li * l2r;
                lr = nr;
                li = ni;
                E = project(E, ux, uy, uz);
                E = cmul_comp(E, load_w(w, bb));
            }
        }
        C3 t = scale(E, lr, li);
        float v[6] = {t.xr, t.xi, t.yr, t.yi, t.zr, t.zi};
        acc.add(v);
    }
    double *o = out + 6 * seg_slot[g];
#pragma unroll
    for (int i = 0; i < 6; ++i)
        o[i] = (double)acc.s[i] - (double)acc.c[i];
}

// Zero fill as a kernel (stream-capturable, unlike a memset node).
__global__ void k_fill_zero(float *__restrict__ p, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * FB + threadIdx.x;
    if (i < n)
        p[i] = 0.0f;
}

// Chunk count for the source-summed mode: about 2^21 (sink, chunk) threads
// (8,192 workgroups, ~4.5 rounds of the chip's resident slots) while keeping
// >= 64 sources per thread.  The mask lets whole waves skip sources, so the
// work per workgroup is uneven and small workgroups balance it (the 64x64
// conformal array forward: 0.66 ms with 2^19 threads, 0.575 ms with 2^21).
// (sink, chunk) threads per launch: 2^LEM_LOS_THREADS_LOG2
#ifndef LEM_LOS_THREADS_LOG2
#define LEM_LOS_THREADS_LOG2 21
#endif

inline int64_t los_chunks(int64_t S, int64_t K)
{
    int64_t want = ceil_div(1ll << LEM_LOS_THREADS_LOG2, K > 0 ? K : 1);
    int64_t cap = ceil_div(S, 64);
    int64_t nc = want < cap ? want : cap;
    return nc < 1 ? 1 : nc;
}

int64_t los_chunk_count(int64_t S, int64_t K) { return los_chunks(S, K); }

int launch_los_finish(const float *ws_sum, const float *ws_cmp, int64_t NC, int64_t K,
                      const float *snk_n, const float *snk_w, int tangent, float *out,
                      hipStream_t st)
{
    hipLaunchKernelGGL(k_field_los_finish, dim3((unsigned)ceil_div(K, FB)), dim3(FB), 0, st,
                       ws_sum, ws_cmp, NC, K, snk_n, snk_w, tangent, out);
    return check_launch("k_field_los_finish");
}

// Projected-area pattern of visiblespace (rayfunctions.py:316-427 with
// angle_pop / patternsort, :431-468): for each shell point k, the sum over the
// sources s with an unblocked ray s -> k of area[s] * cos(angle between the
// ray and the source normal), angles beyond 90 degrees contributing nothing.
// The reference's |cos(+-arccos(dp))| of the clamped dot product dp of the unit
// vectors is dp itself; area == nullptr means a unit portion (its
// vertex_area == 0 case).  One thread per shell point, sources in index order,
// float64 sum.
__global__ __launch_bounds__(256) void k_projected_area(
    const float *__restrict__ src, const float *__restrict__ nrm,
    const float *__restrict__ area, int64_t S, const float *__restrict__ snk, int64_t K,
    const unsigned long long *__restrict__ bits, float *__restrict__ out)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= K)
        return;
    const int64_t W = (K + 63) / 64;
    const unsigned long long bit = 1ull << (k & 63);
    const float tx = snk[3 * k], ty = snk[3 * k + 1], tz = snk[3 * k + 2];
    double acc = 0.0;
    for (int64_t s = 0; s < S; ++s) {
        if (bits[s * W + (k >> 6)] & bit)
            continue;  // blocked
        const Ray r = charge(src[3 * s], src[3 * s + 1], src[3 * s + 2], tx, ty, tz);
        const float nx = nrm[3 * s], ny = nrm[3 * s + 1], nz = nrm[3 * s + 2];
        const float nn = sqrtf(nx * nx + ny * ny + nz * nz);
        const float rn = sqrtf(r.dx * r.dx + r.dy * r.dy + r.dz * r.dz);
        float dp = (r.dx / rn) * (nx / nn) + (r.dy / rn) * (ny / nn);
        dp = dp + (r.dz / rn) * (nz / nn);
        dp = fminf(fmaxf(dp, -1.0f), 1.0f);
        if (dp > 0.0f)
            acc += area ? (double)area[s] * (double)dp : 1.0;
    }
    out[k] = (float)acc;
}

// Source-summed field with NC source chunks whose partials go to ws
// (2 * NC * 6 * K floats: Kahan sums, then compensations).  Every partial is
// written before the merge reads it, so ws needs no clearing.
static int field_los_sum(const float *src, const float *src_w, int64_t S, const float *snk,
                         const float *snk_n, const float *snk_w, int64_t K,
                         const unsigned long long *hb, int64_t w64, int sm, double alpha,
                         double beta, int tangent, float *out, int64_t NC, float *ws,
                         hipStream_t st)
{
    const int64_t nbk = ceil_div(K, FB);
    const int64_t CH = ceil_div(S, NC);
    float *ws_sum = ws;
    float *ws_cmp = ws_sum + NC * 6 * K;
    LEM_REQUIRE(NC * nbk <= 2147483647LL, "lem_field_los: grid too large");
    if (alpha != 0.0)
        hipLaunchKernelGGL(k_field_los_partial<true>, dim3((unsigned)(NC * nbk)), dim3(FB), 0, st,
                           src, src_w, S, snk, snk_n, K, hb, w64, sm, (float)alpha, beta, CH, nbk,
                           ws_sum, ws_cmp);
    else
        hipLaunchKernelGGL(k_field_los_partial<false>, dim3((unsigned)(NC * nbk)), dim3(FB), 0, st,
                           src, src_w, S, snk, snk_n, K, hb, w64, sm, (float)alpha, beta, CH, nbk,
                           ws_sum, ws_cmp);
    int rc = check_launch("k_field_los_partial");
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_field_los_finish, dim3((unsigned)nbk), dim3(FB), 0, st, ws_sum, ws_cmp,
                       NC, K, snk_n, snk_w, tangent, out);
    return check_launch("k_field_los_finish");
}

}  // namespace lem

using namespace lem;

extern "C" int64_t lem_field_los_workspace(int64_t S, int64_t K, int flags)
{
    if (flags & LEM_FIELD_ELEMENTS)
        return 0;
    if (S <= 0 || K <= 0)
        return 0;
    int64_t nc = los_chunks(S, K);
    return 2 * nc * 6 * K * (int64_t)sizeof(float);
}

extern "C" int lem_field_los(const float *src, const float *src_w, int64_t S, const float *snk,
                             const float *snk_n, const float *snk_w, int64_t K,
                             const uint64_t *hit_bits, double alpha, double beta, int flags,
                             float *out, void *workspace, int64_t workspace_bytes, void *stream)
{
    clear_error();
    LEM_REQUIRE(S >= 0 && K >= 0, "lem_field_los: negative size");
    if (K == 0)
        return OK;
    LEM_REQUIRE(out && snk && snk_n, "lem_field_los: null pointer");
    LEM_REQUIRE(S == 0 || (src && src_w), "lem_field_los: null source arrays");
    const int sm = (flags & LEM_FIELD_BITS_SINK_MAJOR) ? 1 : 0;
    const int64_t w64 = sm ? ceil_div(S, 64) : ceil_div(K, 64);
    const int tangent = (flags & LEM_FIELD_SINK_TANGENT) ? 1 : 0;
    hipStream_t st = as_stream(stream);
    const unsigned long long *hb = reinterpret_cast<const unsigned long long *>(hit_bits);
    const int64_t nbk = ceil_div(K, FB);
    if (flags & LEM_FIELD_ELEMENTS) {
        if (S == 0)
            return OK;
        LEM_REQUIRE(S * nbk <= 2147483647LL, "lem_field_los: S*K too large for one launch");
        if (alpha != 0.0)
            hipLaunchKernelGGL(k_field_los_elements<true>, dim3((unsigned)(S * nbk)), dim3(FB), 0,
                               st, src, src_w, snk, snk_n, snk_w, K, hb, w64, sm, (float)alpha,
                               beta, tangent, nbk, out);
        else
            hipLaunchKernelGGL(k_field_los_elements<false>, dim3((unsigned)(S * nbk)), dim3(FB), 0,
                               st, src, src_w, snk, snk_n, snk_w, K, hb, w64, sm, (float)alpha,
                               beta, tangent, nbk, out);
        return check_launch("k_field_los_elements");
    }
    if (S == 0) {
        hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)ceil_div(6 * K, FB)), dim3(FB), 0, st, out,
                           6 * K);
        return check_launch("k_fill_zero");
    }
    const int64_t need = lem_field_los_workspace(S, K, flags);
    if (!workspace || workspace_bytes < need) {
        set_error("lem_field_los: workspace smaller than lem_field_los_workspace()");
        return E_WORKSPACE;
    }
    return field_los_sum(src, src_w, S, snk, snk_n, snk_w, K, hb, w64, sm, alpha, beta, tangent,
                         out, los_chunks(S, K), reinterpret_cast<float *>(workspace), st);
}

extern "C" int lem_field_paths(const float *pos, const float *nrm, const float *w,
                               const float *eta, int64_t N,
                               const int32_t *index, int ncol, const int64_t *order,
                               const int64_t *seg_start, const int64_t *seg_slot, int64_t nseg,
                               double alpha, double beta, double *out, int *err_flag,
                               void *stream)
{
    clear_error();
    LEM_REQUIRE(N >= 0 && nseg >= 0, "lem_field_paths: negative size");
    LEM_REQUIRE(ncol >= 2 && ncol <= 4, "lem_field_paths: index must have 2 to 4 columns");
    if (nseg == 0)
        return OK;
    LEM_REQUIRE(pos && nrm && w && index && order && seg_start && seg_slot && out && err_flag,
                "lem_field_paths: null pointer");
    LEM_REQUIRE(ceil_div(nseg, FB) <= 2147483647LL, "lem_field_paths: too many segments");
    if (alpha != 0.0)
        hipLaunchKernelGGL(k_field_paths<true>, dim3((unsigned)ceil_div(nseg, FB)), dim3(FB), 0,
                       as_stream(stream), pos, nrm, w, eta, N, index, ncol, order, seg_start, seg_slot,
                       nseg, (float)alpha, beta, out, err_flag);
    else
        hipLaunchKernelGGL(k_field_paths<false>, dim3((unsigned)ceil_div(nseg, FB)), dim3(FB), 0,
                       as_stream(stream), pos, nrm, w, eta, N, index, ncol, order, seg_start, seg_slot,
                       nseg, (float)alpha, beta, out, err_flag);
    return check_launch("k_field_paths");
}

extern "C" int lem_projected_area(const float *src, const float *nrm, const float *area,
                                  int64_t S, const float *snk, int64_t K, const uint64_t *hit_bits,
                                  float *out, void *stream)
{
    clear_error();
    LEM_REQUIRE(S >= 0 && K >= 0, "lem_projected_area: negative size");
    if (K == 0)
        return 0;
    LEM_REQUIRE(snk && out && (S == 0 || (src && nrm && hit_bits)),
                "lem_projected_area: null pointer");
    hipLaunchKernelGGL(k_projected_area, dim3((unsigned)ceil_div(K, 256)), dim3(256), 0,
                       as_stream(stream), src, nrm, area, S, snk, K,
                       reinterpret_cast<const unsigned long long *>(hit_bits), out);
    return check_launch("k_projected_area");
}

// The device address of `scratch` on the current device, looked up on the
// first call on that device and kept.
static int scratch_on_current_device(float **ptr)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error(std::string("lem_field_los_noalloc: hipGetDevice: ") + hipGetErrorString(e));
        return E_HIP;
    }
    if (dev < 0 || dev >= MAX_DEVICES) {
        set_error("lem_field_los_noalloc: device index out of range");
        return E_INVALID;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("lem_field_los_noalloc: hipGetSymbolAddress: ") +
                      hipGetErrorString(e));
            return E_HIP;
        }
        g_scratch_ptr[dev] = static_cast<float *>(p);
    }
    *ptr = g_scratch_ptr[dev];
    return OK;
}

extern "C" int64_t lem_field_los_noalloc_max_sinks(void)
{
    return NOALLOC_SCRATCH_FLOATS / 12;
}

extern "C" int lem_field_los_noalloc(const float *src, const float *src_w, int64_t S,
                                     const float *snk, const float *snk_n, const float *snk_w,
                                     int64_t K, const uint64_t *hit_bits, double alpha,
                                     double beta, int flags, float *out, void *stream)
{
    clear_error();
    LEM_REQUIRE(S >= 0 && K >= 0, "lem_field_los_noalloc: negative size");
    if (K == 0)
        return OK;
    // the per-element network and the empty source set use no scratch
    if ((flags & LEM_FIELD_ELEMENTS) || S == 0)
        return lem_field_los(src, src_w, S, snk, snk_n, snk_w, K, hit_bits, alpha, beta, flags,
                             out, nullptr, 0, stream);
    LEM_REQUIRE(out && snk && snk_n, "lem_field_los_noalloc: null pointer");
    LEM_REQUIRE(src && src_w, "lem_field_los_noalloc: null source arrays");
    if (K > NOALLOC_SCRATCH_FLOATS / 12) {
        set_error("lem_field_los_noalloc: K = " + std::to_string(K) + " sinks exceed the " +
                  std::to_string(NOALLOC_SCRATCH_FLOATS / 12) + " that its " +
                  std::to_string(LEM_FIELD_LOS_NOALLOC_SCRATCH_BYTES) +
                  "-byte static scratch holds; use lem_field_los with a workspace");
        return E_INVALID;
    }
    float *ws = nullptr;
    int rc = scratch_on_current_device(&ws);
    if (rc)
        return rc;
    // as many source chunks as lem_field_los uses, fewer if their partials
    // would not fit the scratch (then the sums are grouped differently)
    int64_t NC = los_chunks(S, K);
    const int64_t fit = NOALLOC_SCRATCH_FLOATS / (12 * K);
    if (NC > fit)
        NC = fit;
    const int sm = (flags & LEM_FIELD_BITS_SINK_MAJOR) ? 1 : 0;
    const int64_t w64 = sm ? ceil_div(S, 64) : ceil_div(K, 64);
    return field_los_sum(src, src_w, S, snk, snk_n, snk_w, K,
                         reinterpret_cast<const unsigned long long *>(hit_bits), w64, sm, alpha,
                         beta, (flags & LEM_FIELD_SINK_TANGENT) ? 1 : 0, out, NC, ws,
                         as_stream(stream));
}
