// lem_adjoint.hip -- hand-written adjoint of the line-of-sight field (gfx950).
//
// Forward (lem_field_los, source sum):   E_k = sum_s vis(s,k) c_sk,
//   c_sk = F(r) (d . n_k) (I - d d^T) w_s,  d = (p_k - p_s)/r,
//   F(r) = -(1/2pi) (-g - 1/r) e^{-g r}/r,  g = alpha + j beta
// (lossy_propagation, empropagation.py:768-801, and the launch projector,
// :178-249).  Given gE_k = dLoss/dE_k in PyTorch's convention
// (dL/dRe + j dL/dIm) and the frozen mask, this computes
//   grad_w_s = sum_k vis conj(F cos) P(d) gE_k
//   grad_p_s = Re sum_k vis [ -F' cos (U.Q) d - F (U.Q)(n_k - d cos)/r
//                             + F cos ((d.w) M U + (U.d) Q)/r ]
// with U = conj(gE_k), Q = P(d) w_s, M = I - d d^T and
// F' = dF/dr = -(1/2pi) e^{-g r}/r (g^2 + 2g/r + 2/r^2); bilinear dots (no
// conjugation).  Sinks are staged through LDS, lanes run over sources; the
// terms of every 16 sinks are summed in float32, those partials are
// Kahan-accumulated, and chunk partials merge in a fixed order.  The
// sink-weight and sink-tangent steps of the forward are applied to gE on entry
// (their transposes), so the adjoint matches lem_field_los for every flag.
#include "lem_device.h"
#include "lem_field_dev.h"
#include "lem_internal.h"
#include "../../include/lem.h"

namespace lem {

constexpr int AB = 256;      // sources per block
constexpr int AK_TILE = 256; // sinks per LDS tile

struct SnkLds {
    double x, y, z;
    float nx, ny, nz;
    float g[6];  // gE (re, im) x 3
};

// gE <- P(n_k) (conj(w_k) * gE): transposes of the forward's tangent step and
// sink-weight product, in reverse order.
__global__ __launch_bounds__(AB) void k_adjoint_prep(const float *__restrict__ gE,
                                                     const float *__restrict__ snk_n,
                                                     const float *__restrict__ snk_w, int tangent,
                                                     int64_t K, float *__restrict__ g_out)
{
    int64_t k = (int64_t)blockIdx.x * AB + threadIdx.x;
    if (k >= K)
        return;
    C3 g = load_w(gE, k);
    if (tangent)
        g = project_normal(g, snk_n[3 * k], snk_n[3 * k + 1], snk_n[3 * k + 2]);
    if (snk_w) {
        C3 w = load_w(snk_w, k);
        w.xi = -w.xi; w.yi = -w.yi; w.zi = -w.zi;
        g = cmul_comp(g, w);
    }
    float2 *o = reinterpret_cast<float2 *>(g_out) + 3 * k;
    o[0] = make_float2(g.xr, g.xi);
    o[1] = make_float2(g.yr, g.yi);
    o[2] = make_float2(g.zr, g.zi);
}

// Partials: block b covers sources [(b % nbs)*256, +256), sink chunk c = b / nbs.
template <bool LOSSY>
__global__ __launch_bounds__(AB) void k_adjoint_partial(
    const float *__restrict__ src, const float *__restrict__ src_w, int64_t S,
    const float *__restrict__ snk, const float *__restrict__ snk_n, const float *__restrict__ g,
    int64_t K, const unsigned long long *__restrict__ hit_bits, int64_t W64, float alpha,
    double beta, int64_t CH, int64_t nbs, const int32_t *__restrict__ order,
    float *__restrict__ ws_sum, float *__restrict__ ws_cmp)
{
    // a * b + c may be fused in the gradient arithmetic (the Kahan updates
    // have no products, so they are not contracted)
#pragma clang fp contract(fast)
    __shared__ SnkLds s_k[AK_TILE];
    const int64_t c = (int64_t)blockIdx.x / nbs;
    const int64_t slot = ((int64_t)blockIdx.x % nbs) * AB + threadIdx.x;
    const bool active = slot < S;
    // lanes over a spatial order of the sources: neighbouring elements see the
    // same sinks, so the mask skips whole waves more often
    const int64_t s = active ? (order ? (int64_t)order[slot] : slot) : 0;
    const int64_t ss = s;
    const double xs = src[3 * ss], ys = src[3 * ss + 1], zs = src[3 * ss + 2];
    const C3 w = load_w(src_w, ss);
    const int64_t k0 = c * CH;
    const int64_t k1 = (k0 + CH) < K ? (k0 + CH) : K;
    const float beta_f = (float)beta;
    const float beta2 = beta_f * beta_f;
    const float al = LOSSY ? alpha : 0.0f;  // lossless: the alpha terms fold away
    const float FRONT = -0.15915494309189535f;
    KahanC3 gw;
    gw.zero();
    float gp[3] = {0.0f, 0.0f, 0.0f}, gpc[3] = {0.0f, 0.0f, 0.0f};
    for (int64_t base = k0; base < k1; base += AK_TILE) {
        __syncthreads();
        {
            int64_t j = base + threadIdx.x;
            SnkLds L;
            if (j < k1) {
                L.x = snk[3 * j]; L.y = snk[3 * j + 1]; L.z = snk[3 * j + 2];
                L.nx = snk_n[3 * j]; L.ny = snk_n[3 * j + 1]; L.nz = snk_n[3 * j + 2];
                C3 gg = load_w(g, j);
                L.g[0] = gg.xr; L.g[1] = gg.xi; L.g[2] = gg.yr;
                L.g[3] = gg.yi; L.g[4] = gg.zr; L.g[5] = gg.zi;
            } else {
                L.x = L.y = L.z = 0.0;
                L.nx = L.ny = L.nz = 0.0f;
                for (int i = 0; i < 6; ++i)
                    L.g[i] = 0.0f;
            }
            s_k[threadIdx.x] = L;
        }
        __syncthreads();
        const int n = (int)((k1 - base) < AK_TILE ? (k1 - base) : AK_TILE);
        if (!active)
            continue;
        unsigned long long word = 0;
        for (int j0 = 0; j0 < n; j0 += SUM_GROUP) {
            const int j1 = (j0 + SUM_GROUP) < n ? (j0 + SUM_GROUP) : n;
            float h[9] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 2
            for (int j = j0; j < j1; ++j) {
                const int64_t k = base + j;
                if (hit_bits) {
                    if (j == 0 || (k & 63) == 0)
                        word = hit_bits[s * W64 + (k >> 6)];
                    if ((word >> (k & 63)) & 1ull)
                        continue;
                }
                const SnkLds &L = s_k[j];
                double dx = L.x - xs, dy = L.y - ys, dz = L.z - zs;
                Phase p = phase_of(dx, dy, dz, beta);
                const float ir = p.inv_r;
                const float ux = (float)dx * ir, uy = (float)dy * ir, uz = (float)dz * ir;
                const float cs = ux * L.nx + uy * L.ny + uz * L.nz;
                // FRONT G = FRONT e^{-alpha r}/r e^{-j beta r}
                float amp = ir * FRONT;
                if (LOSSY)
                    amp *= expf(-alpha * (float)p.r);
                const float Gr = amp * p.c, Gi = amp * p.s;
                // F = FRONT (-g - 1/r) G ; F' = FRONT (g^2 + 2g/r + 2/r^2) G
                const float ar = -al - ir;
                const float Fr = ar * Gr + beta_f * Gi, Fi = ar * Gi - beta_f * Gr;
                const float br = (al * al - beta2) + 2.0f * ir * (al + ir);
                const float bi = 2.0f * beta_f * (al + ir);
                const float dFr = br * Gr - bi * Gi, dFi = br * Gi + bi * Gr;
                // g = gE_k; bilinear dots (no conjugation) of d, w and U = conj(g)
                const float gxr = L.g[0], gxi = L.g[1], gyr = L.g[2], gyi = L.g[3], gzr = L.g[4],
                            gzi = L.g[5];
                const float uwr = ux * w.xr + uy * w.yr + uz * w.zr;
                const float uwi = ux * w.xi + uy * w.yi + uz * w.zi;
                const float ugr = ux * gxr + uy * gyr + uz * gzr;  // d.g; U.d = (ugr, -ugi)
                const float ugi = ux * gxi + uy * gyi + uz * gzi;
                // P(d) g
                const float Pxr = gxr - ux * ugr, Pxi = gxi - ux * ugi;
                const float Pyr = gyr - uy * ugr, Pyi = gyi - uy * ugi;
                const float Pzr = gzr - uz * ugr, Pzi = gzi - uz * ugi;
                // U.Q with Q = P(d) w: P is real and symmetric, so U.Q = conj(P(d) g).w
                const float UQr = Pxr * w.xr + Pxi * w.xi + Pyr * w.yr + Pyi * w.yi +
                                  Pzr * w.zr + Pzi * w.zi;
                const float UQi = Pxr * w.xi - Pxi * w.xr + Pyr * w.yi - Pyi * w.yr +
                                  Pzr * w.zi - Pzi * w.zr;
                // weight gradient: conj(F cos) P(d) g
                {
                    const float Lr = Fr * cs, Li = Fi * cs;
                    h[0] = h[0] + Lr * Pxr; h[0] = h[0] + Li * Pxi;
                    h[1] = h[1] + Lr * Pxi; h[1] = h[1] - Li * Pxr;
                    h[2] = h[2] + Lr * Pyr; h[2] = h[2] + Li * Pyi;
                    h[3] = h[3] + Lr * Pyi; h[3] = h[3] - Li * Pyr;
                    h[4] = h[4] + Lr * Pzr; h[4] = h[4] + Li * Pzi;
                    h[5] = h[5] + Lr * Pzi; h[5] = h[5] - Li * Pzr;
                }
                // position gradient: with M U = U - d (U.d) and Q = w - d (d.w),
                //   (d.w) M U + (U.d) Q = (d.w) U + (U.d) w - 2 (d.w)(U.d) d
                // so the sum is Re[coef d + a2 n + c1 U + c2 w] with
                //   a1 = -F' cos (U.Q), a2 = -F (U.Q)/r, c1 = F cos (d.w)/r,
                //   c2 = F cos (U.d)/r, coef = a1 - a2 cos - 2 c1 (U.d)
                {
                    const float a1 = -(dFr * UQr - dFi * UQi) * cs;
                    const float a2 = -(Fr * UQr - Fi * UQi) * ir;
                    const float csr = cs * ir;
                    const float Fcr = Fr * csr, Fci = Fi * csr;
                    const float c1r = Fcr * uwr - Fci * uwi, c1i = Fcr * uwi + Fci * uwr;
                    const float c2r = Fcr * ugr + Fci * ugi, c2i = Fci * ugr - Fcr * ugi;
                    const float coef = (a1 - a2 * cs) - 2.0f * (c1r * ugr + c1i * ugi);
                    // Re(c1 U_x) = c1r gxr + c1i gxi ; Re(c2 w_x) = c2r wxr - c2i wxi
                    h[6] = h[6] + coef * ux; h[6] = h[6] + a2 * L.nx;
                    h[6] = h[6] + c1r * gxr; h[6] = h[6] + c1i * gxi;
                    h[6] = h[6] + c2r * w.xr; h[6] = h[6] - c2i * w.xi;
                    h[7] = h[7] + coef * uy; h[7] = h[7] + a2 * L.ny;
                    h[7] = h[7] + c1r * gyr; h[7] = h[7] + c1i * gyi;
                    h[7] = h[7] + c2r * w.yr; h[7] = h[7] - c2i * w.yi;
                    h[8] = h[8] + coef * uz; h[8] = h[8] + a2 * L.nz;
                    h[8] = h[8] + c1r * gzr; h[8] = h[8] + c1i * gzi;
                    h[8] = h[8] + c2r * w.zr; h[8] = h[8] - c2i * w.zi;
                }
            }
            gw.add(h);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                float y = h[6 + i] - gpc[i];
                float t = gp[i] + y;
                gpc[i] = (t - gp[i]) - y;
                gp[i] = t;
            }
        }
    }
    if (active) {
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            ws_sum[(c * 9 + i) * S + s] = gw.s[i];
            ws_cmp[(c * 9 + i) * S + s] = gw.c[i];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            ws_sum[(c * 9 + 6 + i) * S + s] = gp[i];
            ws_cmp[(c * 9 + 6 + i) * S + s] = gpc[i];
        }
    }
}

// Fixed-order merge of the chunk partials: one thread per (source, gradient
// component) -- 9 S threads, each summing its component over the chunks in
// chunk order.
__global__ __launch_bounds__(AB) void k_adjoint_finish(const float *__restrict__ ws_sum,
                                                       const float *__restrict__ ws_cmp,
                                                       int64_t NC, int64_t S,
                                                       float *__restrict__ grad_w,
                                                       float *__restrict__ grad_p)
{
    const int64_t t = (int64_t)blockIdx.x * AB + threadIdx.x;
    if (t >= 9 * S)
        return;
    const int i = (int)(t / S);
    const int64_t s = t - i * S;
    // eight chunks' loads issued together, added in chunk order
    const float *ps = ws_sum + (int64_t)i * S + s, *pc = ws_cmp + (int64_t)i * S + s;
    const int64_t step = 9 * S;
    double a = 0.0;
    int64_t c = 0;
    for (; c + 8 <= NC; c += 8) {
        float vs[8], vc[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            vs[u] = ps[(c + u) * step];
            vc[u] = pc[(c + u) * step];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            a += (double)vs[u] - (double)vc[u];
    }
    for (; c < NC; ++c)
        a += (double)ps[c * step] - (double)pc[c * step];
    if (i < 6) {
        if (grad_w)
            grad_w[6 * s + i] = (float)a;
    } else if (grad_p) {
        grad_p[3 * s + (i - 6)] = (float)a;
    }
}

// Sink chunks: enough (source, chunk) threads to fill the chip, chunk length a
// multiple of the LDS tile (keeps 64-bit mask words aligned).
// (source, chunk) threads per launch: 2^LEM_ADJ_THREADS_LOG2
#ifndef LEM_ADJ_THREADS_LOG2
#define LEM_ADJ_THREADS_LOG2 19
#endif

inline int64_t adj_chunks(int64_t S, int64_t K)
{
    int64_t want = ceil_div(1ll << LEM_ADJ_THREADS_LOG2, S > 0 ? S : 1);
    int64_t cap = ceil_div(K, AK_TILE);
    int64_t nc = want < cap ? want : cap;
    return nc < 1 ? 1 : nc;
}

}  // namespace lem

using namespace lem;

inline int64_t adj_part_bytes(int64_t S, int64_t K)
{
    const int64_t NC = adj_chunks(S, K);
    const int64_t b = 2 * NC * 9 * S * (int64_t)sizeof(float) + 6 * K * (int64_t)sizeof(float);
    return ((b + 255) / 256) * 256;
}

extern "C" int64_t lem_field_los_adjoint_workspace(int64_t S, int64_t K)
{
    if (S <= 0 || K <= 0)
        return 0;
    const int64_t ord = point_order_workspace(S);
    return ord < 0 ? -1 : adj_part_bytes(S, K) + ord;
}

extern "C" int lem_field_los_adjoint(const float *src, const float *src_w, int64_t S,
                                     const float *snk, const float *snk_n, const float *snk_w,
                                     int64_t K, const uint64_t *hit_bits, double alpha,
                                     double beta, int flags, const float *grad_out,
                                     float *grad_w, float *grad_p, void *workspace,
                                     int64_t workspace_bytes, void *stream)
{
    clear_error();
    LEM_REQUIRE(S > 0 && K > 0, "lem_field_los_adjoint: empty source or sink set");
    LEM_REQUIRE(src && src_w && snk && snk_n && grad_out && (grad_w || grad_p),
                "lem_field_los_adjoint: null pointer");
    LEM_REQUIRE(!(flags & LEM_FIELD_ELEMENTS), "lem_field_los_adjoint: source-sum forward only");
    const int64_t need = lem_field_los_adjoint_workspace(S, K);
    if (need < 0) {
        set_error("lem_field_los_adjoint: workspace size query failed");
        return E_HIP;
    }
    if (!workspace || workspace_bytes < need) {
        set_error("lem_field_los_adjoint: workspace smaller than lem_field_los_adjoint_workspace()");
        return E_WORKSPACE;
    }
    const int64_t NC = adj_chunks(S, K);
    const int64_t CH = ceil_div(ceil_div(K, NC), AK_TILE) * AK_TILE;
    const int64_t NCe = ceil_div(K, CH);
    const int64_t nbs = ceil_div(S, AB);
    LEM_REQUIRE(NCe * nbs <= 2147483647LL, "lem_field_los_adjoint: grid too large");
    float *ws_sum = reinterpret_cast<float *>(workspace);
    float *ws_cmp = ws_sum + NC * 9 * S;
    float *g = ws_cmp + NC * 9 * S;
    hipStream_t st = as_stream(stream);
    const int32_t *order = nullptr;
    if (!(flags & LEM_FIELD_INPUT_ORDER)) {
        int rc = point_order(src, S, reinterpret_cast<char *>(workspace) + adj_part_bytes(S, K),
                             &order, st);
        if (rc)
            return rc;
    }
    hipLaunchKernelGGL(k_adjoint_prep, dim3((unsigned)ceil_div(K, AB)), dim3(AB), 0, st, grad_out,
                       snk_n, snk_w, (flags & LEM_FIELD_SINK_TANGENT) ? 1 : 0, K, g);
    if (alpha != 0.0)
        hipLaunchKernelGGL(k_adjoint_partial<true>, dim3((unsigned)(NCe * nbs)), dim3(AB), 0, st, src, src_w,
                       S, snk, snk_n, g, K, reinterpret_cast<const unsigned long long *>(hit_bits),
                       ceil_div(K, 64), (float)alpha, beta, CH, nbs, order, ws_sum, ws_cmp);
    else
        hipLaunchKernelGGL(k_adjoint_partial<false>, dim3((unsigned)(NCe * nbs)), dim3(AB), 0, st, src, src_w,
                       S, snk, snk_n, g, K, reinterpret_cast<const unsigned long long *>(hit_bits),
                       ceil_div(K, 64), (float)alpha, beta, CH, nbs, order, ws_sum, ws_cmp);
    int rc = check_launch("k_adjoint_partial");
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_adjoint_finish, dim3((unsigned)ceil_div(9 * S, AB)), dim3(AB), 0, st,
                       ws_sum, ws_cmp, NCe, S, grad_w, grad_p);
    return check_launch("k_adjoint_finish");
}
