// lem_gw.hip -- batched-excitation contraction E = G . W on FP32 MFMA (gfx950).
//
// For many excitation sets at once (the reference evaluates one excitation per
// calculate_scattering call; its multiE branch cannot run,
// frequency_domain.py:612-617), the source sum of the line-of-sight field is a
// complex matrix product per frequency f and field component c:
//
//   E[f,c,k,b] = sum_s G_c[k,s,f] W[f,s,b]
//   G_c[k,s,f] = F_f(r) (d . n_k) (e - d (d . e))_c          (all paths visible)
//
// with the lossy_propagation factor F_f (empropagation.py:768-801) at
// wavenumber k_f, d = (p_k - p_s)/r and a fixed source polarisation e.
// Each workgroup owns 32 sinks x 64 weight sets at one frequency and walks the
// sources in chunks of 32: the 256 lanes generate the 32 x 32 x 3 complex G
// tile into LDS (float64 path length and phase, as in lem_fields.hip), stage
// the matching 32 x 64 complex W tile, then four wavefronts accumulate with
// v_mfma_f32_32x32x2f32 -- wave w owns weight-set columns 32*(w&1).. and the
// real (w<2) or imaginary (w>=2) part of all three components:
//   Re E += Gr Wr + (-Gi) Wi,   Im E += Gr Wi + Gi Wr.
#include "lem_device.h"
#include "lem_internal.h"
#include "../../include/lem.h"

namespace lem {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int GW_M = 32;   // sinks per workgroup
constexpr int GW_N = 64;   // weight sets (columns)
constexpr int GW_KC = 32;  // sources per chunk

template <bool LOSSY>
__global__ __launch_bounds__(256) void k_gw_mfma(
    const float *__restrict__ src, int64_t S, const float *__restrict__ snk,
    const float *__restrict__ snk_n, int64_t K, float ex, float ey, float ez,
    const double *__restrict__ kf, float alpha, const float2 *__restrict__ W,
    float *__restrict__ out)
{
    __shared__ float sGr[3][GW_KC][GW_M];
    __shared__ float sGi[3][GW_KC][GW_M];
    __shared__ float sWr[GW_KC][GW_N];
    __shared__ float sWi[GW_KC][GW_N];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int ct = wave & 1, part = wave >> 1;
    const int64_t f = blockIdx.y;
    const int64_t k0 = (int64_t)blockIdx.x * GW_M;
    const double kw = kf[f];
    const float kw_f = (float)kw;
    const float FRONT = -0.15915494309189535f;

    // this lane's sink for G generation
    const int kl = tid & 31;
    const int64_t k = k0 + kl;
    const double xk = snk[3 * k], yk = snk[3 * k + 1], zk = snk[3 * k + 2];
    const float nx = snk_n[3 * k], ny = snk_n[3 * k + 1], nz = snk_n[3 * k + 2];

    f32x16 acc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int j = 0; j < 16; ++j)
            acc[c][j] = 0.0f;

    for (int64_t s0 = 0; s0 < S; s0 += GW_KC) {
        __syncthreads();
        // --- G tile: 32 sources x 32 sinks, 4 pairs per lane, one at a time:
        // 78 VGPRs + 48 accumulation registers fit 4 waves per SIMD (unrolled,
        // 97 + 48 allowed 3; 2.6 % slower)
#pragma unroll 1
        for (int i = 0; i < 4; ++i) {
            const int sl = (tid >> 5) + 8 * i;
            const int64_t s = s0 + sl;
            double dx = xk - (double)src[3 * s], dy = yk - (double)src[3 * s + 1],
                   dz = zk - (double)src[3 * s + 2];
            Phase p = phase_of(dx, dy, dz, kw);
            const float ir = p.inv_r;
            const float ux = (float)dx * ir, uy = (float)dy * ir, uz = (float)dz * ir;
            const float cs = ux * nx + uy * ny + uz * nz;
            float amp = ir;
            if (LOSSY)
                amp *= expf(-alpha * (float)p.r);
            const float Gr = amp * p.c, Gi = amp * p.s;
            const float ar = -alpha - ir, ai = -kw_f;
            const float f_ = FRONT * cs;
            const float Lr = (ar * Gr - ai * Gi) * f_;
            const float Li = (ar * Gi + ai * Gr) * f_;
            const float ue = ux * ex + uy * ey + uz * ez;
            const float px = ex - ux * ue, py = ey - uy * ue, pz = ez - uz * ue;
            sGr[0][sl][kl] = Lr * px; sGi[0][sl][kl] = Li * px;
            sGr[1][sl][kl] = Lr * py; sGi[1][sl][kl] = Li * py;
            sGr[2][sl][kl] = Lr * pz; sGi[2][sl][kl] = Li * pz;
        }
        // --- W tile: 32 sources x 64 sets
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int idx = tid + 256 * i;
            const int sl = idx >> 6, b = idx & 63;
            float2 w = W[(f * S + s0 + sl) * GW_N + b];
            sWr[sl][b] = w.x;
            sWi[sl][b] = w.y;
        }
        __syncthreads();
        // --- MFMA: 16 k-steps of 2 sources
#pragma unroll 4
        for (int kk = 0; kk < GW_KC / 2; ++kk) {
            const int si = 2 * kk + (lane >> 5);
            const int row = lane & 31;
            const int col = ct * 32 + (lane & 31);
            const float wr = sWr[si][col], wi = sWi[si][col];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float gr = sGr[c][si][row], gi = sGi[c][si][row];
                if (part == 0) {
                    acc[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(gr, wr, acc[c], 0, 0, 0);
                    acc[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(-gi, wi, acc[c], 0, 0, 0);
                } else {
                    acc[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(gr, wi, acc[c], 0, 0, 0);
                    acc[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(gi, wr, acc[c], 0, 0, 0);
                }
            }
        }
    }
    // --- D[i][j]: lane holds column j = lane % 32, register r holds row
    //     i = 8*(r/4) + 4*(lane/32) + r%4
    const int j = ct * 32 + (lane & 31);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int64_t kk2 = k0 + i;
            out[((((f * 3 + c) * K) + kk2) * GW_N + j) * 2 + part] = acc[c][r];
        }
    }
}

}  // namespace lem

using namespace lem;

extern "C" int lem_gw_mfma(const float *src, int64_t S, float pol_x, float pol_y, float pol_z,
                           const float *snk, const float *snk_n, int64_t K,
                           const double *wavenumbers, int64_t F, double alpha, const float *W,
                           int64_t B, float *out, void *stream)
{
    clear_error();
    LEM_REQUIRE(S > 0 && K > 0 && F > 0, "lem_gw_mfma: empty problem");
    LEM_REQUIRE(src && snk && snk_n && wavenumbers && W && out, "lem_gw_mfma: null pointer");
    LEM_REQUIRE(B == GW_N, "lem_gw_mfma: B must be 64 weight sets");
    LEM_REQUIRE(S % GW_KC == 0, "lem_gw_mfma: S must be a multiple of 32 (pad with zero weights)");
    LEM_REQUIRE(K % GW_M == 0, "lem_gw_mfma: K must be a multiple of 32 (pad the sinks)");
    LEM_REQUIRE(F <= 65535 && K / GW_M <= 2147483647LL, "lem_gw_mfma: grid too large");
    dim3 grid((unsigned)(K / GW_M), (unsigned)F);
    if (alpha != 0.0)
        hipLaunchKernelGGL(k_gw_mfma<true>, grid, dim3(256), 0, as_stream(stream), src, S, snk, snk_n, K,
                       pol_x, pol_y, pol_z, wavenumbers, (float)alpha,
                       reinterpret_cast<const float2 *>(W), out);
    else
        hipLaunchKernelGGL(k_gw_mfma<false>, grid, dim3(256), 0, as_stream(stream), src, S, snk, snk_n, K,
                       pol_x, pol_y, pol_z, wavenumbers, (float)alpha,
                       reinterpret_cast<const float2 *>(W), out);
    return check_launch("k_gw_mfma");
}
