// lem_loss.hip -- the far-field dB mask loss and its gradient in one pass
// (gfx950).
//
// For a source-summed far field E (K x 3 complex64), the power of sink k is
// P_k = sum_c |E_kc|^2 (components added in order), the reference power R is
// the largest P_k (or a given value), and
//   db_k  = 10 log10(P_k / R + eps)
//   err_k = db_k                          inside the main beam
//           max(db_k - sidelobe_db, 0)    outside it
//   loss  = (1/K) sum_k err_k^2.
// R is a constant of the loss (not differentiated), so with PyTorch's
// convention for complex gradients (dL/dRe + j dL/dIm)
//   dloss/dE_kc = (2/K) err_k g_k (10/ln 10) / (P_k + eps R) * 2 E_kc,
// g_k = 1 inside the beam, [db_k > sidelobe_db] outside.  Three launches: the
// largest power (an order-free maximum), the per-sink terms with one partial
// sum per workgroup (a fixed tree), and the sum of the partials (fixed
// strides and tree) -- the loss is deterministic.  No counterpart in the
// reference (it has no gradients); the beam-synthesis objective of the
// differentiable far field.
#include "lem_device.h"
#include "lem_internal.h"
#include "../../include/lem.h"

namespace lem {

constexpr int LB = 256;
constexpr int LOSS_MAX_BLOCKS = 1024;

LEM_DEV float sink_power(const float *__restrict__ E, int64_t k)
{
    const float2 *e = reinterpret_cast<const float2 *>(E) + 3 * k;
    float p = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float2 v = e[c];
        p = p + (v.x * v.x + v.y * v.y);
    }
    return p;
}

// ws[0] <- 0 (the bits of +0.0f); a kernel, not a memset, so that the call can
// be captured in a HIP graph
__global__ void k_loss_init(unsigned int *__restrict__ ws)
{
    if (threadIdx.x == 0)
        ws[0] = 0u;
}

// largest sink power: non-negative floats order like their bit patterns
__global__ __launch_bounds__(LB) void k_loss_peak(const float *__restrict__ E, int64_t K,
                                                  unsigned int *__restrict__ peak_bits)
{
    __shared__ unsigned int s[LB];
    unsigned int m = 0u;
    for (int64_t k = (int64_t)blockIdx.x * LB + threadIdx.x; k < K; k += (int64_t)gridDim.x * LB) {
        const float p = sink_power(E, k);
        const unsigned int b = __float_as_uint(p);
        m = (p == p && b > m) ? b : m;  // a NaN power takes no part
    }
    s[threadIdx.x] = m;
    __syncthreads();
    for (int off = LB / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            s[threadIdx.x] = s[threadIdx.x + off] > s[threadIdx.x] ? s[threadIdx.x + off]
                                                                   : s[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicMax(peak_bits, s[0]);
}

// per-sink error, gradient and one partial sum of err^2 per workgroup
__global__ __launch_bounds__(LB) void k_loss_terms(const float *__restrict__ E,
                                                   const uint8_t *__restrict__ in_beam, int64_t K,
                                                   float sidelobe_db, float ref_power, float eps,
                                                   const unsigned int *__restrict__ peak_bits,
                                                   float *__restrict__ grad,
                                                   float *__restrict__ partial)
{
    __shared__ float s[LB];
    const float R = ref_power > 0.0f ? ref_power : __uint_as_float(peak_bits[0]);
    const float TEN_OVER_LN10 = 4.342944819032518f;
    const float gscale = 4.0f / (float)K;  // (2/K) * 2 E
    float acc = 0.0f;
    for (int64_t k = (int64_t)blockIdx.x * LB + threadIdx.x; k < K; k += (int64_t)gridDim.x * LB) {
        const float p = sink_power(E, k);
        const float q = p / R + eps;
        const float db = 10.0f * log10f(q);
        const bool beam = in_beam[k] != 0;
        const float x = db - sidelobe_db;
        const float err = beam ? db : (x > 0.0f ? x : 0.0f);
        acc = acc + err * err;
        if (grad) {
            const bool live = beam || x > 0.0f;
            // d db / dP = (10/ln 10) / (q R)
            const float c = live ? gscale * err * TEN_OVER_LN10 / (q * R) : 0.0f;
            const float2 *e = reinterpret_cast<const float2 *>(E) + 3 * k;
            float2 *g = reinterpret_cast<float2 *>(grad) + 3 * k;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const float2 v = e[i];
                g[i] = make_float2(c * v.x, c * v.y);
            }
        }
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int off = LB / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partial[blockIdx.x] = s[0];
}

// loss = (sum of the partials) / K in float64: thread t adds partials t,
// t + LB, ... in order, then a fixed tree -- the same order every call
__global__ __launch_bounds__(LB) void k_loss_final(const float *__restrict__ partial, int nb,
                                                   int64_t K, float *__restrict__ loss)
{
    __shared__ double s[LB];
    double a = 0.0;
    for (int b = threadIdx.x; b < nb; b += LB)
        a += (double)partial[b];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int off = LB / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        loss[0] = (float)(s[0] / (double)K);
}

inline int loss_blocks(int64_t K)
{
    const int64_t nb = ceil_div(K > 0 ? K : 1, LB);
    return (int)(nb < LOSS_MAX_BLOCKS ? nb : LOSS_MAX_BLOCKS);
}

}  // namespace lem

using namespace lem;

extern "C" int64_t lem_db_mask_loss_workspace(int64_t K)
{
    if (K < 0)
        return -1;
    return 256 + 4 * (int64_t)loss_blocks(K);
}

extern "C" int lem_db_mask_loss(const float *E, const uint8_t *in_beam, int64_t K,
                                double sidelobe_db, double ref_power, double eps, float *loss,
                                float *grad_E, void *workspace, int64_t workspace_bytes,
                                void *stream)
{
    clear_error();
    LEM_REQUIRE(K > 0, "lem_db_mask_loss: no sinks");
    LEM_REQUIRE(E && in_beam && loss, "lem_db_mask_loss: null pointer");
    LEM_REQUIRE(eps >= 0.0, "lem_db_mask_loss: negative eps");
    if (!workspace || workspace_bytes < lem_db_mask_loss_workspace(K)) {
        set_error("lem_db_mask_loss: workspace smaller than lem_db_mask_loss_workspace()");
        return E_WORKSPACE;
    }
    hipStream_t st = as_stream(stream);
    unsigned int *peak_bits = reinterpret_cast<unsigned int *>(workspace);
    float *partial = reinterpret_cast<float *>(reinterpret_cast<char *>(workspace) + 256);
    const int nb = loss_blocks(K);
    if (!(ref_power > 0.0)) {
        hipLaunchKernelGGL(k_loss_init, dim3(1), dim3(64), 0, st, peak_bits);
        hipLaunchKernelGGL(k_loss_peak, dim3((unsigned)nb), dim3(LB), 0, st, E, K, peak_bits);
    }
    hipLaunchKernelGGL(k_loss_terms, dim3((unsigned)nb), dim3(LB), 0, st, E, in_beam, K,
                       (float)sidelobe_db, ref_power > 0.0 ? (float)ref_power : 0.0f, (float)eps,
                       peak_bits, grad_E, partial);
    hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(LB), 0, st, partial, nb, K, loss);
    return check_launch("k_loss_final");
}
