// lem_raycast.hip -- brute-force Moller-Trumbore visibility for gfx950.
//
// Replaces kernel1Dv3 (rayfunctions.py:212-244) as launched by
// launchRaycaster1Dv3 / chunkingRaycaster1Dv3 (rayfunctions.py:1951-2017,
// 2102-2369), and the native raycast_brute_force (raycasting_brute_force.cuh:68-88).
//
// Layout: one thread per ray, 256 rays per block.  The block streams the
// triangle table through LDS in tiles of 256 (v0, e1, e2 padded to 48 B so a
// wave reads a triangle with three broadcast ds_read_b128).  Every ray walks
// the triangles in table order, so the closest-hit mode reproduces kernel1Dv3's
// shrinking window exactly.  The any-hit mode stops a ray at its first accepted
// hit and stops the block as soon as all its rays are done.
#include "lem_device.h"
#include "lem_internal.h"
#include "../../include/lem.h"

namespace lem {

constexpr int RC_BLOCK = 256;
constexpr int RC_TILE = 256;

struct alignas(16) TriLds {
    float v0x, v0y, v0z, e1x;
    float e1y, e1z, e2x, e2y;
    float e2z, pad0, pad1, pad2;
};

__device__ __forceinline__ void stage_tile(TriLds *s_tri, const float *__restrict__ tris,
                                           int64_t base, int64_t T)
{
    int64_t j = base + threadIdx.x;
    TriLds L;
    if (j < T) {
        Tri t = make_tri(tris + 9 * j);
        L.v0x = t.v0x; L.v0y = t.v0y; L.v0z = t.v0z;
        L.e1x = t.e1x; L.e1y = t.e1y; L.e1z = t.e1z;
        L.e2x = t.e2x; L.e2y = t.e2y; L.e2z = t.e2z;
    } else {
        L.v0x = L.v0y = L.v0z = 0.0f;
        L.e1x = L.e1y = L.e1z = 0.0f;
        L.e2x = L.e2y = L.e2z = 0.0f;
    }
    L.pad0 = L.pad1 = L.pad2 = 0.0f;
    s_tri[threadIdx.x] = L;
}

__device__ __forceinline__ Tri load_tri(const TriLds *s_tri, int j)
{
    TriLds L = s_tri[j];
    Tri t;
    t.v0x = L.v0x; t.v0y = L.v0y; t.v0z = L.v0z;
    t.e1x = L.e1x; t.e1y = L.e1y; t.e1z = L.e1z;
    t.e2x = L.e2x; t.e2y = L.e2y; t.e2z = L.e2z;
    return t;
}

// Walk every triangle (or until the first hit with ANY) for one ray.
// SEM 0 = numba predicate + kernel1Dv3 dist update, SEM 1 = native culled.
template <int SEM, bool ANY>
__device__ __forceinline__ void tile_walk(const TriLds *s_tri, int n, const Ray &r,
                                          float &dist, bool &hit)
{
    for (int j = 0; j < n; ++j) {
        Tri t = load_tri(s_tri, j);
        if (SEM == 0) {
            // hit_numba accepts t < dist - 2e-7 only, so an accepted hit is
            // always closer than dist (kernel1Dv3's t < ray.dist test)
            double tt;
            if (hit_numba(r, dist, t, tt)) {
                dist = (float)tt;
                hit = true;
                if (ANY)
                    return;
            }
        } else {
            if (hit_native(r, t)) {
                hit = true;
                if (ANY)
                    return;
            }
        }
    }
}

// Dense S x K: block b owns source s = b / nbk and sinks [(b % nbk)*256, +256).
// hit_bits is S x W64 words; bit (k % 64) of word (k / 64) set = ray blocked.
template <int SEM, bool ANY>
__global__ __launch_bounds__(RC_BLOCK) void k_visibility_dense(
    const float *__restrict__ src, const float *__restrict__ snk, int64_t K,
    const float *__restrict__ tris, int64_t T, int skip_self,
    unsigned long long *__restrict__ hit_bits, int64_t W64, float *__restrict__ dist_out,
    int64_t nbk)
{
    __shared__ TriLds s_tri[RC_TILE];
    const int64_t s = (int64_t)blockIdx.x / nbk;
    const int64_t k = ((int64_t)blockIdx.x % nbk) * RC_BLOCK + threadIdx.x;
    const bool active = k < K;
    Ray r;
    float dist = 0.0f;
    bool hit = false;
    bool done = !active;
    {
        float ox = src[3 * s], oy = src[3 * s + 1], oz = src[3 * s + 2];
        int64_t kk = active ? k : 0;
        r = charge(ox, oy, oz, snk[3 * kk], snk[3 * kk + 1], snk[3 * kk + 2]);
        dist = r.len;
        if (skip_self && s == k)
            done = true;
    }
    for (int64_t base = 0; base < T; base += RC_TILE) {
        __syncthreads();
        stage_tile(s_tri, tris, base, T);
        __syncthreads();
        int n = (int)((T - base) < RC_TILE ? (T - base) : RC_TILE);
        if (!done) {
            tile_walk<SEM, ANY>(s_tri, n, r, dist, hit);
            if (ANY && hit)
                done = true;
        }
        if (ANY) {
            if (__syncthreads_and(done ? 1 : 0))
                break;
        }
    }
    unsigned long long b = __ballot(active && hit);
    if ((threadIdx.x & 63) == 0) {
        int64_t word = k >> 6;
        if (word < W64)
            hit_bits[s * W64 + word] = b;
    }
    if (dist_out && active)
        dist_out[s * K + k] = dist;
}

// Explicit rays: pair i = (origin point, target point), 0-based rows of the
// unified N x 3 point table.  Outputs kernel1Dv3's `intersect` and `dist`.
template <int SEM, bool ANY>
__global__ __launch_bounds__(RC_BLOCK) void k_raycast_pairs(
    const float *__restrict__ pts, int64_t N, const int32_t *__restrict__ pairs, int64_t R,
    const float *__restrict__ tris, int64_t T, uint8_t *__restrict__ intersect,
    float *__restrict__ dist_out)
{
    __shared__ TriLds s_tri[RC_TILE];
    const int64_t i = (int64_t)blockIdx.x * RC_BLOCK + threadIdx.x;
    const bool active = i < R;
    Ray r;
    float dist = 0.0f;
    bool hit = false;
    bool done = !active;
    bool bad = false;
    {
        int64_t ii = active ? i : 0;
        int64_t a = pairs[2 * ii], b = pairs[2 * ii + 1];
        // Rows outside the point table are reported (intersect = 0xFF), never read.
        if (a < 0 || a >= N || b < 0 || b >= N) {
            bad = true;
            a = 0;
            b = 0;
            done = true;
        }
        r = charge(pts[3 * a], pts[3 * a + 1], pts[3 * a + 2], pts[3 * b], pts[3 * b + 1],
                   pts[3 * b + 2]);
        dist = r.len;
    }
    for (int64_t base = 0; base < T; base += RC_TILE) {
        __syncthreads();
        stage_tile(s_tri, tris, base, T);
        __syncthreads();
        int n = (int)((T - base) < RC_TILE ? (T - base) : RC_TILE);
        if (!done) {
            tile_walk<SEM, ANY>(s_tri, n, r, dist, hit);
            if (ANY && hit)
                done = true;
        }
        if (ANY) {
            if (__syncthreads_and(done ? 1 : 0))
                break;
        }
    }
    if (active) {
        intersect[i] = bad ? 0xFF : (hit ? 1 : 0);
        if (dist_out)
            dist_out[i] = dist;
    }
}

}  // namespace lem

using namespace lem;

extern "C" int lem_visibility_dense(const float *src, int64_t S, const float *snk, int64_t K,
                                    const float *tris, int64_t T, int semantics, int flags,
                                    uint64_t *hit_bits, float *dist, void *stream)
{
    clear_error();
    LEM_REQUIRE(S >= 0 && K >= 0 && T >= 0, "lem_visibility_dense: negative size");
    LEM_REQUIRE(semantics == LEM_SEM_NUMBA || semantics == LEM_SEM_NATIVE,
                "lem_visibility_dense: unknown semantics");
    if (S == 0 || K == 0)
        return OK;
    LEM_REQUIRE(src && snk && hit_bits, "lem_visibility_dense: null pointer");
    LEM_REQUIRE(T == 0 || tris, "lem_visibility_dense: null triangle table");
    const bool any = (flags & LEM_VIS_ANY_HIT) != 0;
    const int skip_self = (flags & LEM_VIS_SKIP_SELF) ? 1 : 0;
    LEM_REQUIRE(!(any && dist && semantics == LEM_SEM_NUMBA),
                "lem_visibility_dense: dist needs the closest-hit mode (no LEM_VIS_ANY_HIT)");
    const int64_t W64 = ceil_div(K, 64);
    const int64_t nbk = ceil_div(K, RC_BLOCK);
    LEM_REQUIRE(S * nbk <= 2147483647LL, "lem_visibility_dense: S*K too large for one launch");
    dim3 grid((unsigned)(S * nbk));
    hipStream_t st = as_stream(stream);
    unsigned long long *hb = reinterpret_cast<unsigned long long *>(hit_bits);
    if (semantics == LEM_SEM_NUMBA) {
        if (any)
            hipLaunchKernelGGL((k_visibility_dense<0, true>), grid, dim3(RC_BLOCK), 0, st, src,
                               snk, K, tris, T, skip_self, hb, W64, dist, nbk);
        else
            hipLaunchKernelGGL((k_visibility_dense<0, false>), grid, dim3(RC_BLOCK), 0, st, src,
                               snk, K, tris, T, skip_self, hb, W64, dist, nbk);
    } else {
        hipLaunchKernelGGL((k_visibility_dense<1, true>), grid, dim3(RC_BLOCK), 0, st, src, snk,
                           K, tris, T, skip_self, hb, W64, dist, nbk);
    }
    return check_launch("k_visibility_dense");
}

extern "C" int lem_raycast_pairs(const float *points, int64_t n_points, const int32_t *pairs,
                                 int64_t n_rays, const float *tris, int64_t T, int semantics,
                                 int flags, uint8_t *intersect, float *dist, void *stream)
{
    clear_error();
    LEM_REQUIRE(n_points >= 0 && n_rays >= 0 && T >= 0, "lem_raycast_pairs: negative size");
    LEM_REQUIRE(semantics == LEM_SEM_NUMBA || semantics == LEM_SEM_NATIVE,
                "lem_raycast_pairs: unknown semantics");
    if (n_rays == 0)
        return OK;
    LEM_REQUIRE(points && pairs && intersect, "lem_raycast_pairs: null pointer");
    LEM_REQUIRE(n_points > 0, "lem_raycast_pairs: empty point table");
    LEM_REQUIRE(T == 0 || tris, "lem_raycast_pairs: null triangle table");
    const bool any = (flags & LEM_VIS_ANY_HIT) != 0;
    LEM_REQUIRE(!(any && dist && semantics == LEM_SEM_NUMBA),
                "lem_raycast_pairs: dist needs the closest-hit mode (no LEM_VIS_ANY_HIT)");
    LEM_REQUIRE(ceil_div(n_rays, RC_BLOCK) <= 2147483647LL, "lem_raycast_pairs: too many rays");
    dim3 grid((unsigned)ceil_div(n_rays, RC_BLOCK));
    hipStream_t st = as_stream(stream);
    if (semantics == LEM_SEM_NUMBA) {
        if (any)
            hipLaunchKernelGGL((k_raycast_pairs<0, true>), grid, dim3(RC_BLOCK), 0, st, points,
                               n_points, pairs, n_rays, tris, T, intersect, dist);
        else
            hipLaunchKernelGGL((k_raycast_pairs<0, false>), grid, dim3(RC_BLOCK), 0, st, points,
                               n_points, pairs, n_rays, tris, T, intersect, dist);
    } else {
        hipLaunchKernelGGL((k_raycast_pairs<1, true>), grid, dim3(RC_BLOCK), 0, st, points, n_points,
                           pairs, n_rays, tris, T, intersect, dist);
    }
    return check_launch("k_raycast_pairs");
}
