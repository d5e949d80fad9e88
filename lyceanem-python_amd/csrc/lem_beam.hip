// lem_beam.hip -- beamformed-pattern power of many weight sets on FP32 MFMA
// (gfx950), the hot loop of MaximumDirectivityMap
// (electromagnetics/beamforming.py:484-673).
//
// For weight rows b (command angle x weight scheme) and element patterns
// Et, Ep (elements x pattern points), the beamformed patterns are
//   Bt[b,p] = sum_e W[b,e] Et[e,p],   Bp[b,p] = sum_e W[b,e] Ep[e,p]
// (the reference's np.sum(weights.reshape(n,1,1) * E, axis=0) per command).
// directivity_transform (beamforming.py:1037-1129) then needs, per row, the
// quadrature-weighted power sum_p q[p] (|Bt|^2 + |Bp|^2) and the power at the
// row's command point -- so the patterns are reduced in the epilogue and never
// stored.  Each workgroup owns 32 rows x 64 points and walks the elements in
// chunks of 32: W and both pattern tiles are staged in LDS, four wavefronts
// accumulate with v_mfma_f32_32x32x2f32 (wave w: point columns 32*(w&1).., the
// real (w<2) or imaginary (w>=2) part, both polarisations), then re and im
// meet in LDS for the power epilogue.  Per-tile partial sums are merged in a
// fixed order by k_beam_finish.
#include "lem_internal.h"
#include "../../include/lem.h"

#include <hip/hip_runtime.h>

namespace lem {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int BM = 32;  // weight rows per workgroup
constexpr int BN = 64;  // pattern points per workgroup
constexpr int BK = 32;  // elements per chunk

__global__ __launch_bounds__(256) void k_beam_power(
    const float2 *__restrict__ W, int64_t ne, const float2 *__restrict__ Et,
    const float2 *__restrict__ Ep, int64_t np, const float *__restrict__ q,
    const int32_t *__restrict__ pstar, float *__restrict__ part, float *__restrict__ ustar,
    int64_t nb)
{
    // chunk staging (40 KB); reused as the epilogue exchange (32 KB)
    __shared__ float smem[2 * BK * BM + 4 * BK * BN];
    // W tile [e][b ^ e]: the row index is XOR-swizzled by the element so that
    // the transposing stores (32 lanes of one row, consecutive elements) and
    // the MFMA operand reads (consecutive rows of one element) both fall in
    // distinct LDS banks
    float *sWr = smem, *sWi = smem + BK * BM;
    float *sE = smem + 2 * BK * BM;                          // [pol][re/im][e][p]
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int ct = wave & 1, part_ri = wave >> 1;
    const int64_t b0 = (int64_t)blockIdx.x * BM;
    const int64_t tile = blockIdx.y;
    const int64_t p0 = tile * BN;

    f32x16 at, ap;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        at[j] = 0.0f;
        ap[j] = 0.0f;
    }
    for (int64_t e0 = 0; e0 < ne; e0 += BK) {
        __syncthreads();
        // W tile: 32 rows x 32 elements, stored [e][b ^ e]
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int idx = tid + 256 * i;
            const int bl = idx >> 5, el = idx & 31;
            const float2 w = W[(b0 + bl) * ne + e0 + el];
            sWr[el * BM + (bl ^ el)] = w.x;
            sWi[el * BM + (bl ^ el)] = w.y;
        }
        // pattern tiles: 32 elements x 64 points, both polarisations
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int idx = tid + 256 * i;
            const int el = idx >> 6, pl = idx & 63;
            const int64_t g = (e0 + el) * np + p0 + pl;
            const float2 t = Et[g], f = Ep[g];
            sE[(0 * BK + el) * BN + pl] = t.x;
            sE[(1 * BK + el) * BN + pl] = t.y;
            sE[(2 * BK + el) * BN + pl] = f.x;
            sE[(3 * BK + el) * BN + pl] = f.y;
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < BK / 2; ++kk) {
            const int e = 2 * kk + (lane >> 5);
            const int row = lane & 31;
            const int col = ct * 32 + (lane & 31);
            const float wr = sWr[e * BM + (row ^ e)], wi = sWi[e * BM + (row ^ e)];
            const float tr = sE[(0 * BK + e) * BN + col], ti = sE[(1 * BK + e) * BN + col];
            const float fr = sE[(2 * BK + e) * BN + col], fi = sE[(3 * BK + e) * BN + col];
            if (part_ri == 0) {  // Re = wr er - wi ei
                at = __builtin_amdgcn_mfma_f32_32x32x2f32(wr, tr, at, 0, 0, 0);
                at = __builtin_amdgcn_mfma_f32_32x32x2f32(-wi, ti, at, 0, 0, 0);
                ap = __builtin_amdgcn_mfma_f32_32x32x2f32(wr, fr, ap, 0, 0, 0);
                ap = __builtin_amdgcn_mfma_f32_32x32x2f32(-wi, fi, ap, 0, 0, 0);
            } else {             // Im = wr ei + wi er
                at = __builtin_amdgcn_mfma_f32_32x32x2f32(wr, ti, at, 0, 0, 0);
                at = __builtin_amdgcn_mfma_f32_32x32x2f32(wi, tr, at, 0, 0, 0);
                ap = __builtin_amdgcn_mfma_f32_32x32x2f32(wr, fi, ap, 0, 0, 0);
                ap = __builtin_amdgcn_mfma_f32_32x32x2f32(wi, fr, ap, 0, 0, 0);
            }
        }
    }
    // D[i][j]: lane holds column j = lane % 32, register r holds row
    // i = 8*(r/4) + 4*(lane/32) + r%4.  Exchange through LDS as
    // X[pol][re/im][i][j] over the 64 columns.
    __syncthreads();
    float *X = smem;  // 4 * 32 * 64 floats
    {
        const int j = ct * 32 + (lane & 31);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            X[((0 * 2 + part_ri) * BM + i) * BN + j] = at[r];
            X[((1 * 2 + part_ri) * BM + i) * BN + j] = ap[r];
        }
    }
    __syncthreads();
    // thread t: row i = t / 8, columns 8*(t%8) .. +7
    const int i = tid >> 3, c0 = (tid & 7) * 8;
    const int64_t b = b0 + i;
    const int32_t ps = pstar[b];
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int j = c0 + c;
        const float tr = X[((0) * BM + i) * BN + j], ti = X[((1) * BM + i) * BN + j];
        const float fr = X[((2) * BM + i) * BN + j], fi = X[((3) * BM + i) * BN + j];
        const float ut = tr * tr + ti * ti, up = fr * fr + fi * fi;
        acc += q[p0 + j] * (ut + up);
        if (ps == p0 + j) {
            ustar[2 * b] = ut;
            ustar[2 * b + 1] = up;
        }
    }
    // the 8 threads of a row are consecutive lanes
#pragma unroll
    for (int off = 4; off > 0; off >>= 1)
        acc += __shfl_down(acc, off, 8);
    if ((tid & 7) == 0)
        part[tile * nb + b] = acc;
}

// power[b] = sum over point tiles, in tile order, float64
__global__ __launch_bounds__(256) void k_beam_finish(const float *__restrict__ part,
                                                     int64_t ntiles, int64_t nb,
                                                     double *__restrict__ power)
{
    const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (b >= nb)
        return;
    double s = 0.0;
    for (int64_t t = 0; t < ntiles; ++t)
        s += (double)part[t * nb + b];
    power[b] = s;
}

}  // namespace lem

using namespace lem;

extern "C" int64_t lem_beam_workspace(int64_t nb, int64_t np)
{
    if (nb < 0 || np < 0)
        return -1;
    return 4 * nb * ((np + BN - 1) / BN);
}

extern "C" int lem_beam_power(const float *W, int64_t nb, const float *Et, const float *Ep,
                              int64_t ne, int64_t np, const float *q, const int32_t *pstar,
                              double *power, float *ustar, void *workspace,
                              int64_t workspace_bytes, void *stream)
{
    clear_error();
    LEM_REQUIRE(nb > 0 && ne > 0 && np > 0, "lem_beam_power: empty problem");
    LEM_REQUIRE(nb % BM == 0, "lem_beam_power: weight rows must be a multiple of 32 (pad)");
    LEM_REQUIRE(ne % BK == 0, "lem_beam_power: elements must be a multiple of 32 (zero-pad)");
    LEM_REQUIRE(np % BN == 0, "lem_beam_power: points must be a multiple of 64 (zero-pad q)");
    LEM_REQUIRE(W && Et && Ep && q && pstar && power && ustar, "lem_beam_power: null pointer");
    LEM_REQUIRE(nb / BM <= 2147483647LL && np / BN <= 65535, "lem_beam_power: grid too large");
    if (!workspace || workspace_bytes < lem_beam_workspace(nb, np)) {
        set_error("lem_beam_power: workspace smaller than lem_beam_workspace()");
        return E_WORKSPACE;
    }
    hipStream_t st = as_stream(stream);
    const int64_t ntiles = np / BN;
    float *part = reinterpret_cast<float *>(workspace);
    dim3 grid((unsigned)(nb / BM), (unsigned)ntiles);
    hipLaunchKernelGGL(k_beam_power, grid, dim3(256), 0, st,
                       reinterpret_cast<const float2 *>(W), ne,
                       reinterpret_cast<const float2 *>(Et), reinterpret_cast<const float2 *>(Ep),
                       np, q, pstar, part, ustar, nb);
    int rc = check_launch("k_beam_power");
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_beam_finish, dim3((unsigned)ceil_div(nb, 256)), dim3(256), 0, st, part,
                       ntiles, nb, power);
    return check_launch("k_beam_finish");
}
