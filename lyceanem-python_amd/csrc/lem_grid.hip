// lem_grid.hip -- uniform-grid acceleration structure and DDA visibility (gfx950).
//
// Build (replaces bin_counts_to_numpy / bin_triangles_to_numpy,
// terrain_acceleration_build.cuh:287-431, whose offsets are an O(B^2) loop in a
// single 512-thread block):
//   plan : scene bounds (device reduction), resolution on the host, per-cell
//          reference counts (each triangle's box inflated by eps, so every cell
//          it touches holds it), exclusive scan -> total references.
//   fill : slots by atomic cursor, then each cell's list sorted by triangle id
//          (deterministic order), then the (v0, e1, e2) records gathered in cell
//          order so a cell's triangles are one contiguous 48-byte-stride run.
// Traversal (replaces the tile "DDA" of raycasting_accelerated.cuh:106-268,
// which scans every tile of the bounding rectangle): Amanatides-Woo stepping
// over the cells the ray segment crosses, boundary distances recomputed from the
// ray origin at every step (no accumulated error), and the numba predicate of
// lem_device.h on every triangle of each visited cell.  The closest-hit mode
// keeps kernel1Dv3's shrinking window and stops once the best hit lies before
// the current cell's exit; the any-hit mode stops at the first accepted hit.
// Masks are identical to brute force: triangle boxes are inflated by eps, which
// exceeds the float32 error of the cell stepping, so a triangle the predicate
// accepts always sits in a visited cell.
#include "lem_device.h"
#include "lem_field_dev.h"
#include "lem_grid_walk.h"
#include "lem_internal.h"
#include "../../include/lem.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstring>

#include <hipcub/hipcub.hpp>

namespace lem {

constexpr int GB = 256;

// Order-preserving float <-> uint key (for atomic min / max of floats).
LEM_DEV unsigned int fkey(float f)
{
    unsigned int u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
inline float fkey_inv(unsigned int k)
{
    unsigned int u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

__global__ void k_bounds_init(unsigned int *b)
{
    if (threadIdx.x < 3)
        b[threadIdx.x] = 0xffffffffu;  // min keys
    else if (threadIdx.x < 6)
        b[threadIdx.x] = 0u;  // max keys
}

__global__ __launch_bounds__(GB) void k_bounds(const float *__restrict__ tris, int64_t n_vals,
                                               unsigned int *__restrict__ b)
{
    __shared__ unsigned int smin[3], smax[3];
    if (threadIdx.x < 3) {
        smin[threadIdx.x] = 0xffffffffu;
        smax[threadIdx.x] = 0u;
    }
    __syncthreads();
    unsigned int lmin[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu};
    unsigned int lmax[3] = {0u, 0u, 0u};
    // n_vals = 3 * number of vertices; value i has axis i % 3
    for (int64_t v = (int64_t)blockIdx.x * GB + threadIdx.x; v * 3 < n_vals;
         v += (int64_t)gridDim.x * GB) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            unsigned int k = fkey(tris[3 * v + a]);
            lmin[a] = k < lmin[a] ? k : lmin[a];
            lmax[a] = k > lmax[a] ? k : lmax[a];
        }
    }
    // wave minima / maxima first, then one LDS atomic per wave and bound
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned int m0 = __shfl_xor(lmin[a], off), m1 = __shfl_xor(lmax[a], off);
            lmin[a] = m0 < lmin[a] ? m0 : lmin[a];
            lmax[a] = m1 > lmax[a] ? m1 : lmax[a];
        }
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            atomicMin(&smin[a], lmin[a]);
            atomicMax(&smax[a], lmax[a]);
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        atomicMin(&b[threadIdx.x], smin[threadIdx.x]);
        atomicMax(&b[3 + threadIdx.x], smax[threadIdx.x]);
    }
}

__global__ __launch_bounds__(GB) void k_count(const float *__restrict__ tris, int64_t T, GridDev G,
                                              int *__restrict__ counts)
{
    int64_t t = (int64_t)blockIdx.x * GB + threadIdx.x;
    if (t >= T)
        return;
    int lo[3], hi[3];
    tri_cell_range(G, tris + 9 * t, lo, hi);
    for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
            for (int x = lo[0]; x <= hi[0]; ++x)
                atomicAdd(&counts[((int64_t)z * G.ny + y) * G.nx + x], 1);
}

// Exclusive scan, three phases: per-block scan + block sums, scan of the block
// sums in one block, add-back.  counts[n] receives the total.
__global__ __launch_bounds__(GB) void k_scan_blocks(const int *__restrict__ in, int64_t n,
                                                    long long *__restrict__ out,
                                                    long long *__restrict__ bsum)
{
    __shared__ long long s[GB];
    int64_t i = (int64_t)blockIdx.x * GB + threadIdx.x;
    long long v = i < n ? (long long)in[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < GB; off <<= 1) {
        long long a = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += a;
        __syncthreads();
    }
    if (i < n)
        out[i] = s[threadIdx.x] - v;
    if (threadIdx.x == GB - 1)
        bsum[blockIdx.x] = s[GB - 1];
}

__global__ __launch_bounds__(1024) void k_scan_sums(long long *__restrict__ bsum, int64_t nb,
                                                    long long *__restrict__ total)
{
    __shared__ long long s[1024];
    // each thread owns a contiguous run of ceil(nb / 1024) sums
    int64_t per = (nb + 1023) / 1024;
    int64_t b0 = threadIdx.x * per;
    int64_t b1 = b0 + per < nb ? b0 + per : nb;
    long long acc = 0;
    for (int64_t b = b0; b < b1; ++b)
        acc += bsum[b];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        long long a = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += a;
        __syncthreads();
    }
    long long run = s[threadIdx.x] - acc;
    for (int64_t b = b0; b < b1; ++b) {
        long long v = bsum[b];
        bsum[b] = run;
        run += v;
    }
    if (threadIdx.x == 1023)
        *total = s[1023];
}

__global__ __launch_bounds__(GB) void k_scan_add(long long *__restrict__ out, int64_t n,
                                                 const long long *__restrict__ bsum)
{
    int64_t i = (int64_t)blockIdx.x * GB + threadIdx.x;
    if (i < n)
        out[i] += bsum[blockIdx.x];
}

__global__ __launch_bounds__(GB) void k_starts(const long long *__restrict__ scan, int64_t n,
                                               long long total, int32_t *__restrict__ cell_start,
                                               int *__restrict__ cursor)
{
    int64_t i = (int64_t)blockIdx.x * GB + threadIdx.x;
    if (i < n) {
        cell_start[i] = (int32_t)scan[i];
        cursor[i] = (int)scan[i];
    } else if (i == n) {
        cell_start[n] = (int32_t)total;
    }
}

__global__ __launch_bounds__(GB) void k_fill(const float *__restrict__ tris, int64_t T, GridDev G,
                                             int *__restrict__ cursor, int32_t *__restrict__ ids)
{
    int64_t t = (int64_t)blockIdx.x * GB + threadIdx.x;
    if (t >= T)
        return;
    int lo[3], hi[3];
    tri_cell_range(G, tris + 9 * t, lo, hi);
    for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
            for (int x = lo[0]; x <= hi[0]; ++x) {
                int slot = atomicAdd(&cursor[((int64_t)z * G.ny + y) * G.nx + x], 1);
                ids[slot] = (int32_t)t;
            }
}

// Each cell's ids in increasing triangle order (shell sort; lists are short).
__global__ __launch_bounds__(GB) void k_sort_cells(const int32_t *__restrict__ cell_start,
                                                   int64_t ncells, int32_t *__restrict__ ids)
{
    int64_t c = (int64_t)blockIdx.x * GB + threadIdx.x;
    if (c >= ncells)
        return;
    int32_t b = cell_start[c], n = cell_start[c + 1] - b;
    int32_t *a = ids + b;
    for (int gap = n / 2; gap > 0; gap /= 2) {
        for (int i = gap; i < n; ++i) {
            int32_t v = a[i];
            int j = i;
            while (j >= gap && a[j - gap] > v) {
                a[j] = a[j - gap];
                j -= gap;
            }
            a[j] = v;
        }
    }
}

__global__ __launch_bounds__(GB) void k_gather(const float *__restrict__ tris,
                                               const int32_t *__restrict__ ids, int64_t nrefs,
                                               float4 *__restrict__ recs)
{
    int64_t i = (int64_t)blockIdx.x * GB + threadIdx.x;
    if (i >= nrefs)
        return;
    float4 r[3];
    make_rec(tris, ids[i], r);
    recs[3 * i + 0] = r[0];
    recs[3 * i + 1] = r[1];
    recs[3 * i + 2] = r[2];
}

// Cell boxes (two float4 per cell: lo.xyz + start, hi.xyz + end).
__global__ __launch_bounds__(GB) void k_cell_boxes(const float *__restrict__ tris,
                                                   const int32_t *__restrict__ cell_start,
                                                   const int32_t *__restrict__ ids, GridDev G,
                                                   int64_t ncells, float4 *__restrict__ box)
{
    int64_t c = (int64_t)blockIdx.x * GB + threadIdx.x;
    if (c >= ncells)
        return;
    cell_box_of(tris, cell_start, ids, G, c, box[2 * c], box[2 * c + 1]);
}

// Macro-block boxes, written after the ncells cell boxes.
__global__ __launch_bounds__(GB) void k_macro_boxes(GridDev G, float4 *__restrict__ box)
{
    const int64_t m = (int64_t)blockIdx.x * GB + threadIdx.x;
    const int64_t nm = (int64_t)G.mx * G.my * G.mz;
    if (m >= nm)
        return;
    float4 lo, hi;
    macro_box_of(G, box, m, lo, hi);
    box[2 * (G.ncells + m)] = lo;
    box[2 * (G.ncells + m) + 1] = hi;
}

// Quantised macro table (8 bytes a block).
__global__ __launch_bounds__(GB) void k_macro_quant(GridDev G, const float4 *__restrict__ box,
                                                    uint2 *__restrict__ mq)
{
    const int64_t m = (int64_t)blockIdx.x * GB + threadIdx.x;
    const int64_t nm = (int64_t)G.mx * G.my * G.mz;
    if (m >= nm)
        return;
    mq[m] = macro_quant_of(G, box, m);
}

// Dense S x K through the grid.
// SINK_MAJOR = false: lane = sink, block = (source, 256 sinks); bits S x W64(K).
// SINK_MAJOR = true : lane = source, block = (sink, 256 sources); bits K x W64(S),
//                     dist K x S.  Neighbouring sources to one sink give nearly
//                     parallel rays, so a wavefront walks nearly the same cells.
// Hardware dispatches workgroup b to XCD b % 8, each XCD with its own L2.
// Logical block L = (b % 8) * (nb / 8) + b / 8 gives every XCD a contiguous run
// of logical blocks -- neighbouring sources to one sink, walking nearly the
// same cells -- so the cell data each L2 holds is reused by its own waves.
LEM_DEV int64_t xcd_block(int64_t b, int64_t nb)
{
    const int64_t per = nb / 8;
    if (b >= per * 8)
        return b;
    return (b % 8) * per + b / 8;
}

// REV (any-hit): each ray is walked from its sink end (see dda_cast's rt).
template <bool ANY, bool SINK_MAJOR, int SEM, bool REV = false>
__global__ __launch_bounds__(GB) void k_visibility_grid(
    const float *__restrict__ src, int64_t S, const float *__restrict__ snk, int64_t K, GridDev G,
    const int32_t *__restrict__ cell_start, const float4 *__restrict__ recs,
    const float4 *__restrict__ box, const int32_t *__restrict__ order,
    unsigned long long *__restrict__ bits, float *__restrict__ dist_out, int64_t nbl, int skip_self)
{
    const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int64_t outer = blk / nbl;
    const int64_t slot = (blk % nbl) * GB + threadIdx.x;
    const int64_t NI = SINK_MAJOR ? S : K;
    const bool active = slot < NI;
    const int64_t inner = (order && active) ? (int64_t)order[slot] : slot;
    const int64_t s = SINK_MAJOR ? (active ? inner : 0) : outer;
    const int64_t k = SINK_MAJOR ? outer : (active ? inner : 0);
    Ray r = charge(src[3 * s], src[3 * s + 1], src[3 * s + 2], snk[3 * k], snk[3 * k + 1],
                   snk[3 * k + 2]);
    float dist = r.len;
    bool hit = false;
    if (active && !(skip_self && s == k)) {
        if (REV) {
            const Ray rb = charge(snk[3 * k], snk[3 * k + 1], snk[3 * k + 2], src[3 * s],
                                  src[3 * s + 1], src[3 * s + 2]);
            float db = rb.len;
            hit = dda_cast<ANY, SEM>(rb, db, G, cell_start, recs, box, nullptr, nullptr, nullptr,
                                     MacroTab{nullptr, nullptr, nullptr}, &r);
        } else {
            hit = dda_cast<ANY, SEM>(r, dist, G, cell_start, recs, box);
        }
    }
    const int64_t W = (NI + 63) / 64;
    if (order) {
        if (active && hit)
            atomicOr(&bits[outer * W + (inner >> 6)], 1ull << (inner & 63));
    } else {
        unsigned long long bw = __ballot(active && hit);
        if ((threadIdx.x & 63) == 0) {
            int64_t word = slot >> 6;
            if (word < W)
                bits[outer * W + word] = bw;
        }
    }
    if (dist_out && active)
        dist_out[outer * NI + inner] = dist;
}

constexpr int LDS_WG_PER_CU = 6;  // 6 x 25.5 KB of the 160 KB LDS of a CU
// Any-hit walks whose macro table has at most MACRO_LDS_MAX7 blocks run 7
// workgroups per CU (7 x 21.8 KB of LDS, 72 registers): the C3 one-bounce
// masks gain 4 %.  The closest-hit walk keeps 6: at 72 registers it spills
// 14 and loses 6 %.
constexpr int MACRO_LDS_MAX7 = 3712;
constexpr int lds_table(int wpc) { return wpc >= 7 ? MACRO_LDS_MAX7 : MACRO_LDS_MAX; }

// Rays of the persistent traversal with a zero direction component (along a
// grid plane), queued by k_visibility_grid_lds as (outer, lane slot): walked
// here with the general slab test (the macro table in LDS, as there).  A
// result adds the ray's bit (atomic OR) and overwrites its distance.  If the
// queue overflowed (more than par_cap such rays), every ray is rechecked
// instead (a scene whose points share a coordinate plane makes them all such).
template <bool ANY, bool SINK_MAJOR, int SEM, bool REV>
__global__ __launch_bounds__(GB, LDS_WG_PER_CU) void k_visibility_grid_par(
    const float *__restrict__ src, int64_t S, const float *__restrict__ snk, int64_t K, GridDev G,
    const int32_t *__restrict__ cell_start, const float4 *__restrict__ recs,
    const float4 *__restrict__ box, const uint2 *__restrict__ mq, int nmacro,
    const int32_t *__restrict__ order, unsigned long long *__restrict__ bits,
    float *__restrict__ dist_out, int skip_self, const unsigned long long *__restrict__ par_count,
    const uint2 *__restrict__ par_queue, int64_t par_cap, int perm)
{
    const unsigned long long n = *par_count;
    if (n == 0)
        return;  // nothing queued (the usual case)
    __shared__ uint32_t smqa[MACRO_LDS_MAX];
    __shared__ uint16_t smqb[MACRO_LDS_MAX];
    __shared__ float4 saxis[3];
    for (int i = threadIdx.x; i < nmacro; i += GB) {
        const uint2 q = mq[i];
        smqa[i] = q.x;
        smqb[i] = (uint16_t)q.y;
    }
    if (threadIdx.x < 3)
        saxis[threadIdx.x] = axis_entry(G, threadIdx.x);
    __syncthreads();
    const MacroTab smq{smqa, smqb, saxis};
    const int64_t NI = SINK_MAJOR ? S : K, NO = SINK_MAJOR ? K : S;
    const int64_t W = (NI + 63) / 64;
    const bool scan = n > (unsigned long long)par_cap;
    const int64_t total = scan ? NO * NI : (int64_t)n;
    const int64_t stride = (int64_t)gridDim.x * GB;
    for (int64_t i = (int64_t)blockIdx.x * GB + threadIdx.x; i < total; i += stride) {
        int64_t outer, slot;
        if (scan) {
            outer = i / NI;
            slot = i % NI;
        } else {
            const uint2 e = par_queue[i];
            outer = e.x;
            slot = e.y;
        }
        const int64_t inner = order ? (int64_t)order[slot] : slot;
        const int64_t s = SINK_MAJOR ? inner : outer;
        const int64_t k = SINK_MAJOR ? outer : inner;
        if (skip_self && s == k)
            continue;
        const Ray r = charge(src[3 * s], src[3 * s + 1], src[3 * s + 2], snk[3 * k],
                             snk[3 * k + 1], snk[3 * k + 2]);
        const Ray rw = REV ? charge(snk[3 * k], snk[3 * k + 1], snk[3 * k + 2], src[3 * s],
                                    src[3 * s + 1], src[3 * s + 2])
                           : r;
        if (!(rw.dx == 0.0f || rw.dy == 0.0f || rw.dz == 0.0f))
            continue;  // (scan) walked by the persistent kernel
        float dist = r.len, dw = rw.len;
        float &dd = REV ? dw : dist;
        const bool hit = dda_cast<ANY, SEM, true, true>(rw, dd, G, cell_start, recs, box, nullptr,
                                                        nullptr, nullptr, smq, REV ? &r : nullptr);
        // perm: the rows are in lane-slot order until k_unpermute_rows
        const int64_t at = perm ? slot : inner;
        if (hit)
            atomicOr(&bits[outer * W + (at >> 6)], 1ull << (at & 63));
        if (dist_out)
            dist_out[outer * NI + inner] = dist;
    }
}

// Queue of k_visibility_grid_par: a 64-bit count (256 bytes reserved) and
// PAR_QUEUE_CAP (outer, slot) entries.
constexpr int64_t PAR_QUEUE_CAP = 1 << 18;
constexpr int64_t PAR_QUEUE_BYTES = 256 + 8 * PAR_QUEUE_CAP;

// The 8 task counters sit 4 KB apart: device-scope atomics on one cache line
// serialise where the line lives, which capped claiming near 82 M per second.
constexpr int COUNTER_STRIDE = 512;  // uint64 elements
// 64-ray words per claim.  With the rows claimed in a spatial order of the
// outer points, longer claims keep a wave on neighbouring rays longer:
// 3 / 8 / 12 / 16 / 32 words gave C2 (1M-triangle terrain) 1,105 / 1,084 /
// 1,083 / 1,089 / 1,093 ms and C3 (one bounce) 758 / 738 / 736 / 733 / 732 ms.
constexpr int CLAIM_WORDS = 12;

// Persistent form of k_visibility_grid: each workgroup copies the quantised
// macro table (at most lds_table(WPC) blocks) into LDS once; then each
// wavefront repeatedly claims a 64-ray word of output (one atomic on one of
// the 8 counters next_task[0..7]), so waves never wait for one another and
// the load balances whatever number of workgroups is resident.  Dead-block
// tests read LDS.  WPC: workgroups per CU the registers are bounded for.
template <bool ANY, bool SINK_MAJOR, int SEM, bool REV = false, int WPC = LDS_WG_PER_CU>
__global__ __launch_bounds__(GB, WPC) void k_visibility_grid_lds(
    const float *__restrict__ src, int64_t S, const float *__restrict__ snk, int64_t K, GridDev G,
    const int32_t *__restrict__ cell_start, const float4 *__restrict__ recs,
    const float4 *__restrict__ box, const uint2 *__restrict__ mq, int nmacro,
    const int32_t *__restrict__ order, unsigned long long *__restrict__ bits,
    float *__restrict__ dist_out, unsigned long long *__restrict__ next_task, int skip_self,
    unsigned long long *__restrict__ par_count, uint2 *__restrict__ par_queue, int64_t par_cap,
    int perm, const int32_t *__restrict__ outer_order)
{
    __shared__ uint32_t smqa[lds_table(WPC)];
    __shared__ uint16_t smqb[lds_table(WPC)];
    __shared__ float4 saxis[3];
    for (int i = threadIdx.x; i < nmacro; i += GB) {
        const uint2 q = mq[i];
        smqa[i] = q.x;
        smqb[i] = (uint16_t)q.y;
    }
    if (threadIdx.x < 3)
        saxis[threadIdx.x] = axis_entry(G, threadIdx.x);
    __syncthreads();
    const MacroTab smq{smqa, smqb, saxis};
    const int64_t NI = SINK_MAJOR ? S : K, NO = SINK_MAJOR ? K : S;
    const int64_t W = (NI + 63) / 64;
    const int lane = threadIdx.x & 63;
    // Workgroup b runs on XCD b % 8, each XCD with its own L2.  XCD x owns
    // words [x*CH, (x+1)*CH) of every outer row -- a contiguous eighth of the
    // (spatially ordered) lane-side points -- and claims them with its own
    // counter, so its L2 holds the cells of one region; an XCD whose share is
    // done claims from the next counters.
    const int64_t CH = (W + 7) / 8;
    // a claim takes CLAIM_WORDS consecutive words of the owner's eighth
    const int64_t UPR = (CH + CLAIM_WORDS - 1) / CLAIM_WORDS;
    int owner = (int)(blockIdx.x & 7), tried = 0;
    for (;;) {
        unsigned long long t = 0;
        if (lane == 0)
            t = atomicAdd(&next_task[owner * COUNTER_STRIDE], 1ull);
        // the claim is the same for the whole wave: read it into scalar
        // registers, so the claim's indices (and the outer point's loads) stay
        // scalar and the walk keeps its vector registers
        t = ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned int)(t >> 32)) << 32) |
            (unsigned long long)__builtin_amdgcn_readfirstlane((unsigned int)t);
        int64_t oc = (int64_t)(t / (unsigned long long)UPR);
        int64_t u = (int64_t)(t % (unsigned long long)UPR);
        if constexpr (!ANY) {
            // Closest hit: when an XCD's share of a row is long (C2: 171
            // claims), walking it whole for one row sweeps more triangles
            // through the XCD's L2 than it holds; claims then interleave 64
            // spatially neighbouring rows (the index runs over the group's
            // rows first, then along their words), so the waves in flight
            // share a compact patch of lane-side points (C2 +2 %).  Each t
            // below UPR * NO maps to one (row, claim), larger t to oc = NO.
            // The any-hit walks keep one row at a time: their shares are
            // short on the one-bounce legs (1-11 claims), where interleaving
            // measured 0.6 % slower and its divisions cost them registers.
            constexpr int64_t GROUP = 64;
            if (UPR >= GROUP) {
                const int64_t full = (NO / GROUP) * GROUP;
                if (t < (unsigned long long)(UPR * full)) {
                    const int64_t r = (int64_t)(t % (unsigned long long)(UPR * GROUP));
                    u = r / GROUP;
                    oc = (int64_t)(t / (unsigned long long)(UPR * GROUP)) * GROUP + r % GROUP;
                } else {
                    // the last NO - full rows, interleaved the same way
                    const int64_t t2 = (int64_t)(t - (unsigned long long)(UPR * full));
                    const int64_t G2 = NO - full;
                    u = G2 > 0 ? t2 / G2 : 0;
                    oc = (G2 > 0 && u < UPR) ? full + t2 % G2 : NO;
                }
            }
        }
        if (oc >= NO) {
            if (++tried == 8)
                break;
            owner = (owner + 1) & 7;
            continue;
        }
        // claims walk the outer points in a spatial order (consecutive claims
        // share cells in L2); rows are written at the point's own index
        const int64_t outer = outer_order ? (int64_t)outer_order[oc] : oc;
        const int64_t w0 = owner * CH + u * CLAIM_WORDS;
        const int64_t end = (int64_t)(owner + 1) * CH < W ? (int64_t)(owner + 1) * CH : W;
        const int64_t w1 = w0 + CLAIM_WORDS < end ? w0 + CLAIM_WORDS : end;
        for (int64_t word = w0; word < w1; ++word) {
            const int64_t slot = word * 64 + lane;
            const bool active = slot < NI;
            // with an order, lane slot -> the slot-th point of the spatial order
            const int64_t inner = (order && active) ? (int64_t)order[slot] : slot;
            const int64_t s = SINK_MAJOR ? (active ? inner : 0) : outer;
            const int64_t k = SINK_MAJOR ? outer : (active ? inner : 0);
            Ray r = charge(src[3 * s], src[3 * s + 1], src[3 * s + 2], snk[3 * k],
                           snk[3 * k + 1], snk[3 * k + 2]);
            float dist = r.len;
            bool hit = false;
            if (active && !(skip_self && s == k)) {
                // REV: walk from the sink end, test with the charged ray r
                const Ray rw = REV ? charge(snk[3 * k], snk[3 * k + 1], snk[3 * k + 2],
                                            src[3 * s], src[3 * s + 1], src[3 * s + 2])
                                   : r;
                float dw = rw.len;
                float &dd = REV ? dw : dist;
                const Ray *rt = REV ? &r : nullptr;
                // rays along a grid plane need the general slab test: they are
                // queued for k_visibility_grid_par (left unblocked, at full
                // length, here), which keeps that walk's registers and code out
                // of this kernel
                if (rw.dx == 0.0f || rw.dy == 0.0f || rw.dz == 0.0f) {
                    // past the capacity the count only has to stay above it
                    if (__atomic_load_n(par_count, __ATOMIC_RELAXED) <= (unsigned long long)par_cap) {
                        const unsigned long long q = atomicAdd(par_count, 1ull);
                        if (q < (unsigned long long)par_cap)
                            par_queue[q] = make_uint2((unsigned int)outer, (unsigned int)slot);
                    }
                } else {
                    hit = dda_cast<ANY, SEM, true, false>(rw, dd, G, cell_start, recs, box,
                                                          nullptr, nullptr, nullptr, smq, rt);
                }
            }
            if (order && !perm) {
                // the lanes' bits belong to several words (zeroed beforehand)
                if (active && hit)
                    atomicOr(&bits[outer * W + (inner >> 6)], 1ull << (inner & 63));
            } else {
                // one word per wave: in input order, or (perm) in lane-slot
                // order, rewritten by k_unpermute_rows afterwards
                const unsigned long long bw = __ballot(active && hit);
                if (lane == 0)
                    bits[outer * W + word] = bw;
            }
            if (dist_out && active)
                dist_out[outer * NI + inner] = dist;
        }
    }
}

// clears the 8 task counters of the persistent traversal
__global__ void k_zero_u64(unsigned long long *__restrict__ p)
{
    if (threadIdx.x < 8)
        p[threadIdx.x * COUNTER_STRIDE] = 0ull;
}

__global__ __launch_bounds__(GB) void k_zero_words(unsigned long long *__restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * GB + threadIdx.x; i < n; i += (int64_t)gridDim.x * GB)
        p[i] = 0ull;
}

// Spatial order of the points a wavefront's lanes run over: 3-D Morton keys
// on a cubic lattice over the points' bounding box (1024 cells along its
// longest side), sorted stably.  64 consecutive points of this order form a
// compact cluster (an 8 x 8 patch of a planar array) instead of, say, one
// 64-point row, so a wave's rays walk nearly the same cells and their loops
// end together.
LEM_DEV uint32_t spread3(uint32_t v)
{
    // 10 bits -> every third bit of 30
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

LEM_DEV float fkey_dec(unsigned int k)
{
    const unsigned int u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __builtin_bit_cast(float, u);
}

// 30-bit Morton key of point i in the box of the bound keys b (min x, y, z,
// max x, y, z as fkey values)
LEM_DEV uint32_t morton_key(const float *__restrict__ p, int64_t i, const unsigned int *b)
{
    float lo[3], ext = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = fkey_dec(b[a]);
        ext = fmaxf(ext, fkey_dec(b[3 + a]) - lo[a]);
    }
    const float inv = ext > 0.0f ? 1023.0f / ext : 0.0f;
    uint32_t key = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float f = (p[3 * i + a] - lo[a]) * inv;
        const uint32_t q = f >= 1023.0f ? 1023u : (f > 0.0f ? (uint32_t)f : 0u);  // NaN -> 0
        key |= spread3(q) << a;
    }
    return key;
}

__global__ __launch_bounds__(GB) void k_morton_keys(const float *__restrict__ p, int64_t n,
                                                    const unsigned int *__restrict__ b,
                                                    uint32_t *__restrict__ keys,
                                                    int32_t *__restrict__ vals)
{
    const int64_t i = (int64_t)blockIdx.x * GB + threadIdx.x;
    if (i >= n)
        return;
    keys[i] = morton_key(p, i, b);
    vals[i] = (int32_t)i;
}

// Bounds and keys of up to MORTON_SMALL points in one workgroup (one launch
// where k_bounds_init, k_bounds and k_morton_keys take three); the same keys.
constexpr int MORTON_SMALL = 65536;
constexpr int MB = 1024;

__global__ __launch_bounds__(MB) void k_morton_small(const float *__restrict__ p, int64_t n,
                                                     uint32_t *__restrict__ keys,
                                                     int32_t *__restrict__ vals)
{
    __shared__ unsigned int sb[6];
    if (threadIdx.x < 3) {
        sb[threadIdx.x] = 0xffffffffu;
        sb[3 + threadIdx.x] = 0u;
    }
    __syncthreads();
    unsigned int lmin[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu};
    unsigned int lmax[3] = {0u, 0u, 0u};
    for (int64_t i = threadIdx.x; i < n; i += MB) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const unsigned int k = fkey(p[3 * i + a]);
            lmin[a] = k < lmin[a] ? k : lmin[a];
            lmax[a] = k > lmax[a] ? k : lmax[a];
        }
    }
    // wave minima / maxima first: 16 LDS atomics per bound instead of 1024
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned int m0 = __shfl_xor(lmin[a], off), m1 = __shfl_xor(lmax[a], off);
            lmin[a] = m0 < lmin[a] ? m0 : lmin[a];
            lmax[a] = m1 > lmax[a] ? m1 : lmax[a];
        }
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            atomicMin(&sb[a], lmin[a]);
            atomicMax(&sb[3 + a], lmax[a]);
        }
    }
    __syncthreads();
    for (int64_t i = threadIdx.x; i < n; i += MB) {
        keys[i] = morton_key(p, i, sb);
        vals[i] = (int32_t)i;
    }
}

// inv[order[slot]] = slot
__global__ __launch_bounds__(GB) void k_invert_order(const int32_t *__restrict__ order,
                                                     int64_t n, int32_t *__restrict__ inv)
{
    const int64_t slot = (int64_t)blockIdx.x * GB + threadIdx.x;
    if (slot < n)
        inv[order[slot]] = (int32_t)slot;
}

// Rows written in lane-slot order (bit j of word w = the ray of slot 64 w + j)
// rewritten in place in the order of the lane-side points: a workgroup copies
// one row (W words) into LDS, then each wave builds output words by ballot,
// bit i = slot inv[i] of the copy.  Replaces one device-scope atomic OR per
// blocked ray with one plain store per 64 rays.
__global__ __launch_bounds__(GB) void k_unpermute_rows(unsigned long long *__restrict__ bits,
                                                       int64_t NO, int64_t W, int64_t NI,
                                                       const int32_t *__restrict__ inv)
{
    extern __shared__ unsigned long long srow[];  // W words
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int64_t row = blockIdx.x; row < NO; row += gridDim.x) {
        unsigned long long *b = bits + row * W;
        __syncthreads();  // the previous row's reads of srow are done
        for (int64_t w = threadIdx.x; w < W; w += GB)
            srow[w] = b[w];
        __syncthreads();
        for (int64_t w = wave; w < W; w += GB / 64) {
            const int64_t i = w * 64 + lane;
            bool bit = false;
            if (i < NI) {
                const int32_t sl = inv[i];
                bit = (srow[sl >> 6] >> (sl & 63)) & 1ull;
            }
            const unsigned long long bw = __ballot(bit);
            if (lane == 0)
                b[w] = bw;
        }
    }
}

// largest row (in words) k_unpermute_rows stages in LDS
constexpr int64_t UNPERM_MAX_WORDS = 8192;  // 64 KB

// Workspace of lem_visibility_grid: the task counter and point bounds, then
// (unless the input order is kept) keys, values and order of the lane-side
// points and the radix sort's temporary storage.
struct VisWs {
    unsigned long long *counter;
    unsigned int *bounds;
    uint32_t *k0, *k1;
    int32_t *v0, *order;
    void *cub;
    size_t cub_bytes;
};

constexpr int MORTON_BITS = 30;
// workspace head: the 8 task counters (4 KB apart), then the point bounds
constexpr int64_t VIS_HEAD = 8 * 8 * COUNTER_STRIDE + 256;

inline int64_t vis_sort_bytes(int64_t n)
{
    size_t bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t *)nullptr,
                                           (uint32_t *)nullptr, (const int32_t *)nullptr,
                                           (int32_t *)nullptr, (int)n, 0,
                                           MORTON_BITS) != hipSuccess)
        return -1;
    return (int64_t)bytes;
}

inline int64_t vis_layout(int64_t n, size_t cub_bytes, char *base, VisWs *L)
{
    auto al = [](int64_t b) { return ((b + 255) / 256) * 256; };
    int64_t off = VIS_HEAD;
    char *p[4];
    for (int i = 0; i < 4; ++i) {
        p[i] = base ? base + off : nullptr;
        off += al(4 * n);
    }
    if (L) {
        L->counter = reinterpret_cast<unsigned long long *>(base);
        L->bounds = reinterpret_cast<unsigned int *>(base + 8 * 8 * COUNTER_STRIDE);
        L->k0 = reinterpret_cast<uint32_t *>(p[0]);
        L->k1 = reinterpret_cast<uint32_t *>(p[1]);
        L->v0 = reinterpret_cast<int32_t *>(p[2]);
        L->order = reinterpret_cast<int32_t *>(p[3]);
        L->cub = base + off;
        L->cub_bytes = cub_bytes;
    }
    return off + (int64_t)cub_bytes;
}

// L.order = the Morton order of n points (the workspace laid out by vis_layout).
int morton_order(const float *pts, int64_t n, const VisWs &L, hipStream_t st)
{
    if (n <= MORTON_SMALL) {
        hipLaunchKernelGGL(k_morton_small, dim3(1), dim3(MB), 0, st, pts, n, L.k0, L.v0);
    } else {
        hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, st, L.bounds);
        hipLaunchKernelGGL(k_bounds, dim3((unsigned)std::min<int64_t>(ceil_div(n, GB), 1024)),
                           dim3(GB), 0, st, pts, 3 * n, L.bounds);
        hipLaunchKernelGGL(k_morton_keys, dim3((unsigned)ceil_div(n, GB)), dim3(GB), 0, st, pts,
                           n, L.bounds, L.k0, L.v0);
    }
    int rc = check_launch("k_morton_keys");
    if (rc)
        return rc;
    size_t tb = L.cub_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(L.cub, tb, L.k0, L.k1, L.v0, L.order, (int)n, 0,
                                           MORTON_BITS, st) != hipSuccess) {
        set_error("radix sort of the Morton keys failed");
        return E_HIP;
    }
    return OK;
}

int64_t point_order_workspace(int64_t n)
{
    if (n <= 0 || n > 2147483647LL)
        return -1;
    const int64_t cub = vis_sort_bytes(n);
    return cub < 0 ? -1 : vis_layout(n, (size_t)cub, nullptr, nullptr);
}

int point_order(const float *pts, int64_t n, void *ws, const int32_t **order, hipStream_t st)
{
    const int64_t cub = vis_sort_bytes(n);
    if (cub < 0) {
        set_error("radix sort size query failed");
        return E_HIP;
    }
    VisWs L;
    vis_layout(n, (size_t)cub, reinterpret_cast<char *>(ws), &L);
    const int rc = morton_order(pts, n, L, st);
    *order = L.order;
    return rc;
}

// Instrumented traversal: counts the Moller-Trumbore tests and cell visits a
// traversal performs (the algorithmic work of the roofline), no outputs.
template <bool ANY>
__global__ __launch_bounds__(GB) void k_grid_stats(
    const float *__restrict__ src, int64_t S, const float *__restrict__ snk, int64_t K, GridDev G,
    const int32_t *__restrict__ cell_start, const float4 *__restrict__ recs,
    const float4 *__restrict__ box, const uint2 *__restrict__ mq, int nmacro,
    unsigned long long *__restrict__ stats, int64_t nbl)
{
    __shared__ uint32_t smqa[MACRO_LDS_MAX];
    __shared__ uint16_t smqb[MACRO_LDS_MAX];
    __shared__ float4 saxis[3];
    const bool lds = nmacro <= MACRO_LDS_MAX;
    if (lds)
        for (int i = threadIdx.x; i < nmacro; i += GB) {
            const uint2 q = mq[i];
            smqa[i] = q.x;
            smqb[i] = (uint16_t)q.y;
        }
    if (threadIdx.x < 3)
        saxis[threadIdx.x] = axis_entry(G, threadIdx.x);
    __syncthreads();
    const MacroTab smq{smqa, smqb, saxis};
    const int64_t k = (int64_t)blockIdx.x / nbl;
    const int64_t s = ((int64_t)blockIdx.x % nbl) * GB + threadIdx.x;
    unsigned int nt = 0, nc = 0, nm = 0, nw = 0;
    bool hit = false;
    if (s < S) {
        Ray r = charge(src[3 * s], src[3 * s + 1], src[3 * s + 2], snk[3 * k], snk[3 * k + 1],
                       snk[3 * k + 2]);
        float dist = r.len;
        const bool par = r.dx == 0.0f || r.dy == 0.0f || r.dz == 0.0f;
        hit = !lds ? dda_cast<ANY>(r, dist, G, cell_start, recs, box, &nt, &nc, &nm)
              : par ? dda_cast<ANY, 0, true, true>(r, dist, G, cell_start, recs, box, &nt, &nc,
                                                   &nm, smq)
                    : dda_cast<ANY, 0, true, false>(r, dist, G, cell_start, recs, box, &nt, &nc,
                                                    &nm, smq);
        nw = conservative_refs(r, dist, G, cell_start);
    }
    // wave sums, one atomic per wave
    unsigned long long t64 = nt, c64 = nc, h64 = hit ? 1ull : 0ull, m64 = nm, w64 = nw;
    for (int off = 32; off > 0; off >>= 1) {
        t64 += __shfl_down(t64, off);
        c64 += __shfl_down(c64, off);
        h64 += __shfl_down(h64, off);
        m64 += __shfl_down(m64, off);
        w64 += __shfl_down(w64, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&stats[0], t64);
        atomicAdd(&stats[1], c64);
        atomicAdd(&stats[2], h64);
        atomicAdd(&stats[3], m64);
        atomicAdd(&stats[4], w64);
    }
}

// Field source sum fused with the grid traversal: each (source, sink) ray is
// cast any-hit through the grid and, when unblocked, its Green's-function term
// is accumulated -- the mask is never stored.  Same partial layout as
// k_field_los_partial (lem_fields.hip); merged by the same finish kernel.
template <bool LOSSY>
__global__ __launch_bounds__(FB) void k_field_los_grid_partial(
    const float *__restrict__ src, const float *__restrict__ src_w, int64_t S,
    const float *__restrict__ snk, const float *__restrict__ snk_n, int64_t K, GridDev G,
    const int32_t *__restrict__ cell_start, const float4 *__restrict__ recs,
    const float4 *__restrict__ box, const int32_t *__restrict__ order, float alpha, double beta,
    int64_t CH, int64_t nbk, float *__restrict__ ws_sum, float *__restrict__ ws_cmp)
{
    __shared__ SrcLds s_src[FS_TILE];
    const int64_t c = (int64_t)blockIdx.x / nbk;
    const int64_t slot = ((int64_t)blockIdx.x % nbk) * FB + threadIdx.x;
    const bool active = slot < K;
    // lanes over a spatial order of the sinks (each sink's sum is unchanged)
    const int64_t k = active ? (order ? (int64_t)order[slot] : slot) : 0;
    const int64_t kk = k;
    const float xkf = snk[3 * kk], ykf = snk[3 * kk + 1], zkf = snk[3 * kk + 2];
    const double xk = xkf, yk = ykf, zk = zkf;
    const float nkx = snk_n[3 * kk], nky = snk_n[3 * kk + 1], nkz = snk_n[3 * kk + 2];
    const int64_t s0 = c * CH;
    const int64_t s1 = (s0 + CH) < S ? (s0 + CH) : S;
    KahanC3 acc;
    acc.zero();
    for (int64_t base = s0; base < s1; base += FS_TILE) {
        __syncthreads();
        stage_sources(s_src, src, src_w, base, s1);
        __syncthreads();
        int n = (int)((s1 - base) < FS_TILE ? (s1 - base) : FS_TILE);
        if (active) {
            // same groups of sources as k_field_los_partial: identical sums
            for (int j0 = 0; j0 < n; j0 += SUM_GROUP) {
                const int j1 = (j0 + SUM_GROUP) < n ? (j0 + SUM_GROUP) : n;
                float g[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
                for (int j = j0; j < j1; ++j) {
                    const SrcLds &L = s_src[j];
                    Ray r = charge((float)L.x, (float)L.y, (float)L.z, xkf, ykf, zkf);
                    float dist = r.len;
                    if (dda_cast<true>(r, dist, G, cell_start, recs, box))
                        continue;
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

struct PlanWs {
    unsigned int *bounds;  // 6
    long long *total;      // 1
    int *counts;           // cap + 1
    long long *scan;       // cap + 1
    long long *bsum;       // ceil((cap + 1) / GB)
};

inline PlanWs plan_layout(void *ws, int64_t cap)
{
    char *p = reinterpret_cast<char *>(ws);
    PlanWs L;
    L.bounds = reinterpret_cast<unsigned int *>(p);
    L.total = reinterpret_cast<long long *>(p + 64);
    L.counts = reinterpret_cast<int *>(p + 128);
    int64_t off = 128 + ((4 * (cap + 1) + 255) / 256) * 256;
    L.scan = reinterpret_cast<long long *>(p + off);
    off += ((8 * (cap + 1) + 255) / 256) * 256;
    L.bsum = reinterpret_cast<long long *>(p + off);
    return L;
}

inline int64_t plan_bytes(int64_t cap)
{
    int64_t off = 128 + ((4 * (cap + 1) + 255) / 256) * 256;
    off += ((8 * (cap + 1) + 255) / 256) * 256;
    off += 8 * (ceil_div(cap + 1, GB) + 1);
    return off;
}

__global__ void k_zero_int(int *__restrict__ p, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * GB + threadIdx.x;
    if (i < n)
        p[i] = 0;
}

}  // namespace lem

using namespace lem;

extern "C" int64_t lem_grid_workspace(int64_t T, float cells_per_tri)
{
    if (T < 0)
        return 0;
    return plan_bytes(plan_cap(T, cells_per_tri));
}

extern "C" int lem_grid_plan(const float *tris, int64_t T, float cells_per_tri,
                             lem_grid_info *info, void *workspace, int64_t workspace_bytes,
                             void *stream)
{
    clear_error();
    LEM_REQUIRE(T > 0 && tris && info, "lem_grid_plan: needs a non-empty triangle table");
    LEM_REQUIRE(T <= 2000000000LL, "lem_grid_plan: too many triangles");
    const int64_t cap = plan_cap(T, cells_per_tri);
    if (!workspace || workspace_bytes < plan_bytes(cap)) {
        set_error("lem_grid_plan: workspace smaller than lem_grid_workspace()");
        return E_WORKSPACE;
    }
    hipStream_t st = as_stream(stream);
    PlanWs L = plan_layout(workspace, cap);
    hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, st, L.bounds);
    int64_t nvals = 9 * T;
    int64_t nblk = std::min<int64_t>(ceil_div(3 * T, GB), 4096);
    hipLaunchKernelGGL(k_bounds, dim3((unsigned)nblk), dim3(GB), 0, st, tris, nvals, L.bounds);
    int rc = check_launch("k_bounds");
    if (rc)
        return rc;
    unsigned int hb[6];
    if (hipMemcpyAsync(hb, L.bounds, sizeof(hb), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        set_error("lem_grid_plan: bounds readback failed");
        return E_HIP;
    }
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = fkey_inv(hb[a]);
        hi[a] = fkey_inv(hb[3 + a]);
    }
    lem_grid_info g;
    std::memset(&g, 0, sizeof(g));
    LEM_REQUIRE(grid_geometry(lo, hi, T, cells_per_tri, &g),
                "lem_grid_plan: non-finite vertex coordinates");
    GridDev G = to_dev(g);
    hipLaunchKernelGGL(k_zero_int, dim3((unsigned)ceil_div(g.ncells + 1, GB)), dim3(GB), 0, st,
                       L.counts, g.ncells + 1);
    hipLaunchKernelGGL(k_count, dim3((unsigned)ceil_div(T, GB)), dim3(GB), 0, st, tris, T, G,
                       L.counts);
    rc = check_launch("k_count");
    if (rc)
        return rc;
    const int64_t nsb = ceil_div(g.ncells, GB);
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nsb), dim3(GB), 0, st, L.counts, g.ncells,
                       L.scan, L.bsum);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, st, L.bsum, nsb, L.total);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nsb), dim3(GB), 0, st, L.scan, g.ncells, L.bsum);
    rc = check_launch("grid scan");
    if (rc)
        return rc;
    long long total = 0;
    if (hipMemcpyAsync(&total, L.total, sizeof(total), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        set_error("lem_grid_plan: total readback failed");
        return E_HIP;
    }
    LEM_REQUIRE(total < 2147483647LL, "lem_grid_plan: more than 2^31 triangle references");
    g.nrefs = total;
    *info = g;
    return OK;
}

extern "C" int lem_grid_fill(const float *tris, int64_t T, const lem_grid_info *info,
                             int32_t *cell_start, int32_t *tri_ids, float *recs, float *cell_box,
                             void *workspace, int64_t workspace_bytes, float cells_per_tri,
                             void *stream)
{
    clear_error();
    LEM_REQUIRE(T > 0 && tris && info && cell_start && tri_ids && recs && cell_box,
                "lem_grid_fill: null argument");
    const int64_t cap = plan_cap(T, cells_per_tri);
    LEM_REQUIRE(info->ncells <= cap, "lem_grid_fill: info does not match the workspace");
    if (!workspace || workspace_bytes < plan_bytes(cap)) {
        set_error("lem_grid_fill: workspace smaller than lem_grid_workspace()");
        return E_WORKSPACE;
    }
    hipStream_t st = as_stream(stream);
    PlanWs L = plan_layout(workspace, cap);
    GridDev G = to_dev(*info);
    const int64_t nc = info->ncells;
    // counts is reused as the insertion cursor
    hipLaunchKernelGGL(k_starts, dim3((unsigned)ceil_div(nc + 1, GB)), dim3(GB), 0, st, L.scan, nc,
                       (long long)info->nrefs, cell_start, L.counts);
    hipLaunchKernelGGL(k_fill, dim3((unsigned)ceil_div(T, GB)), dim3(GB), 0, st, tris, T, G,
                       L.counts, tri_ids);
    hipLaunchKernelGGL(k_sort_cells, dim3((unsigned)ceil_div(nc, GB)), dim3(GB), 0, st, cell_start,
                       nc, tri_ids);
    if (info->nrefs > 0)
        hipLaunchKernelGGL(k_gather, dim3((unsigned)ceil_div(info->nrefs, GB)), dim3(GB), 0, st,
                           tris, tri_ids, info->nrefs, reinterpret_cast<float4 *>(recs));
    hipLaunchKernelGGL(k_cell_boxes, dim3((unsigned)ceil_div(nc, GB)), dim3(GB), 0, st, tris,
                       cell_start, tri_ids, G, nc, reinterpret_cast<float4 *>(cell_box));
    hipLaunchKernelGGL(k_macro_boxes, dim3((unsigned)ceil_div(info->nmacro, GB)), dim3(GB), 0, st,
                       G, reinterpret_cast<float4 *>(cell_box));
    hipLaunchKernelGGL(k_macro_quant, dim3((unsigned)ceil_div(info->nmacro, GB)), dim3(GB), 0, st,
                       G, reinterpret_cast<const float4 *>(cell_box),
                       reinterpret_cast<uint2 *>(cell_box + 8 * (nc + info->nmacro)));
    return check_launch("lem_grid_fill");
}

// Workspace region of the outer points' order (after the lane-side layout):
// keys, values, order and the sort's storage for NO points.
inline int64_t outer_region(int64_t NO, size_t cub_bytes, char *base, VisWs *L)
{
    auto al = [](int64_t b) { return ((b + 255) / 256) * 256; };
    if (L) {
        L->k0 = reinterpret_cast<uint32_t *>(base);
        L->k1 = reinterpret_cast<uint32_t *>(base + al(4 * NO));
        L->v0 = reinterpret_cast<int32_t *>(base + 2 * al(4 * NO));
        L->order = reinterpret_cast<int32_t *>(base + 3 * al(4 * NO));
        L->cub = base + 4 * al(4 * NO);
        L->cub_bytes = cub_bytes;
    }
    return 4 * al(4 * NO) + al((int64_t)cub_bytes);
}

extern "C" int64_t lem_visibility_grid_workspace(int64_t S, int64_t K, int flags)
{
    if (S < 0 || K < 0)
        return -1;
    const int64_t NI = (flags & LEM_VIS_SINK_MAJOR) ? S : K;
    const int64_t NO = (flags & LEM_VIS_SINK_MAJOR) ? K : S;
    int64_t base = VIS_HEAD;
    if (!(flags & LEM_VIS_INPUT_ORDER) && NI > 0) {
        if (NI > 2147483647LL)
            return -1;
        const int64_t cub = vis_sort_bytes(NI);
        if (cub < 0)
            return -1;
        base = ceil_div(vis_layout(NI, (size_t)cub, nullptr, nullptr), 256) * 256;
        if (NO > 0 && NO <= 2147483647LL) {
            const int64_t cubo = vis_sort_bytes(NO);
            if (cubo < 0)
                return -1;
            base += outer_region(NO, (size_t)cubo, nullptr, nullptr);
        }
    }
    // the queue of rays along grid planes (k_visibility_grid_par) comes last
    return ceil_div(base, 256) * 256 + PAR_QUEUE_BYTES;
}

extern "C" int lem_visibility_grid(const float *src, int64_t S, const float *snk, int64_t K,
                                   const lem_grid_info *info, const int32_t *cell_start,
                                   const float *recs, const float *cell_box, int semantics,
                                   int flags, uint64_t *hit_bits, float *dist, void *workspace,
                                   int64_t workspace_bytes, void *stream)
{
    clear_error();
    LEM_REQUIRE(S >= 0 && K >= 0, "lem_visibility_grid: negative size");
    if (S == 0 || K == 0)
        return OK;
    LEM_REQUIRE(src && snk && info && cell_start && recs && cell_box && hit_bits,
                "lem_visibility_grid: null pointer");
    LEM_REQUIRE(info->n[0] > 0 && info->n[1] > 0 && info->n[2] > 0, "lem_visibility_grid: bad grid");
    LEM_REQUIRE(semantics == LEM_SEM_NUMBA || semantics == LEM_SEM_NATIVE,
                "lem_visibility_grid: unknown semantics");
    const int skip_self = (flags & LEM_VIS_SKIP_SELF) ? 1 : 0;
    const bool any = (flags & LEM_VIS_ANY_HIT) != 0 || semantics == LEM_SEM_NATIVE;
    const bool sink_major = (flags & LEM_VIS_SINK_MAJOR) != 0;
    const bool rev = (flags & LEM_VIS_FROM_SINK) != 0;
    LEM_REQUIRE(!(any && dist), "lem_visibility_grid: dist needs the closest-hit mode");
    LEM_REQUIRE(!rev || any, "lem_visibility_grid: LEM_VIS_FROM_SINK needs the any-hit mode");
    const int64_t NI = sink_major ? S : K, NO = sink_major ? K : S;
    LEM_REQUIRE(NI <= 2147483647LL, "lem_visibility_grid: more than 2^31-1 lane-side points");
    LEM_REQUIRE(NO <= 4294967295LL, "lem_visibility_grid: more than 2^32-1 outer points");
    const int64_t nbl = ceil_div(NI, GB);
    GridDev G = to_dev(*info);
    hipStream_t st = as_stream(stream);
    const int64_t need = lem_visibility_grid_workspace(S, K, flags);
    if (need < 0) {
        set_error("lem_visibility_grid: workspace size query failed");
        return E_HIP;
    }
    if (!workspace || workspace_bytes < need) {
        set_error("lem_visibility_grid: workspace smaller than lem_visibility_grid_workspace()");
        return E_WORKSPACE;
    }
    unsigned long long *hb = reinterpret_cast<unsigned long long *>(hit_bits);
    const float4 *rc4 = reinterpret_cast<const float4 *>(recs);
    const float4 *bx4 = reinterpret_cast<const float4 *>(cell_box);
    const uint2 *mq = reinterpret_cast<const uint2 *>(cell_box + 8 * (info->ncells + info->nmacro));
    VisWs L;
    const int64_t cub = (flags & LEM_VIS_INPUT_ORDER) ? 0 : vis_sort_bytes(NI);
    vis_layout(NI, (size_t)std::max<int64_t>(cub, 0), reinterpret_cast<char *>(workspace), &L);
    const int32_t *order = nullptr, *outer_order = nullptr;
    // persistent walk over a spatial order with rows that fit in LDS: each wave
    // stores its 64 bits in lane-slot order and the rows are reordered after
    // the walk (no per-ray atomics, no clearing); otherwise bits are ORed per
    // lane into cleared rows
    const int64_t Wn = ceil_div(NI, 64);
    const bool perm = !(flags & LEM_VIS_INPUT_ORDER) && info->nmacro <= MACRO_LDS_MAX &&
                      Wn <= UNPERM_MAX_WORDS;
    if (!(flags & LEM_VIS_INPUT_ORDER)) {
        // spatial order of the lane-side points
        int rc = morton_order(sink_major ? src : snk, NI, L, st);
        if (rc)
            return rc;
        if (perm) {
            // the sort's value input (L.v0) is free again: the inverse order
            hipLaunchKernelGGL(k_invert_order, dim3((unsigned)ceil_div(NI, GB)), dim3(GB), 0, st,
                               L.order, NI, L.v0);
        } else {
            const int64_t nw = NO * Wn;
            hipLaunchKernelGGL(k_zero_words,
                               dim3((unsigned)std::min<int64_t>(ceil_div(nw, GB), 65536)), dim3(GB),
                               0, st, hb, nw);
        }
        rc = check_launch("k_zero_words / k_invert_order");
        if (rc)
            return rc;
        order = L.order;
        if (info->nmacro <= MACRO_LDS_MAX && NO <= 2147483647LL) {
            // the persistent walk's claims also follow a spatial order of the
            // outer points
            const int64_t cubo = vis_sort_bytes(NO);
            if (cubo < 0) {
                set_error("lem_visibility_grid: radix sort size query failed");
                return E_HIP;
            }
            const int64_t off =
                ceil_div(vis_layout(NI, (size_t)std::max<int64_t>(cub, 0), nullptr, nullptr), 256) *
                256;
            VisWs Lo = L;
            outer_region(NO, (size_t)cubo, reinterpret_cast<char *>(workspace) + off, &Lo);
            rc = morton_order(sink_major ? snk : src, NO, Lo, st);
            if (rc)
                return rc;
            outer_order = Lo.order;
        }
    }
    if (info->nmacro <= MACRO_LDS_MAX) {
        // persistent launch: as many workgroups as are resident at once.  The
        // registers are bounded for wpc waves per SIMD (launch bounds), so the
        // LDS per CU sets the count.
        int dev = 0, cus = 0, lds = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) !=
                hipSuccess) {
            set_error("lem_visibility_grid: device query failed");
            return E_HIP;
        }
        const bool w7 = any && semantics != LEM_SEM_NATIVE && info->nmacro <= MACRO_LDS_MAX7;
        const int wpc = w7 ? 7 : LDS_WG_PER_CU;
        const int per_cu = std::max(
            1, std::min(wpc, lds / (int)(lds_table(wpc) * (sizeof(uint32_t) + sizeof(uint16_t)))));
        const int64_t words = NO * ceil_div(NI, 64);
        const int64_t nwg =
            std::max<int64_t>(1, std::min<int64_t>(ceil_div(words, 4), (int64_t)cus * per_cu));
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, L.counter);
        // queue of the rays along grid planes, walked afterwards by
        // k_visibility_grid_par over as many workgroups as fill the chip
        char *qb = reinterpret_cast<char *>(workspace) + (need - PAR_QUEUE_BYTES);
        unsigned long long *pcount = reinterpret_cast<unsigned long long *>(qb);
        uint2 *pq = reinterpret_cast<uint2 *>(qb + 256);
        hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(GB), 0, st, pcount, 1);
#define LEM_VGL(A, M, SM, R, WP)                                                              \
    do {                                                                                      \
        hipLaunchKernelGGL((k_visibility_grid_lds<A, M, SM, R, WP>), dim3((unsigned)nwg), dim3(GB), 0, \
                           st, src, S, snk, K, G, cell_start, rc4, bx4, mq, info->nmacro, order, \
                           hb, dist, L.counter, skip_self, pcount, pq, PAR_QUEUE_CAP,          \
                           perm ? 1 : 0, outer_order);                                         \
        hipLaunchKernelGGL((k_visibility_grid_par<A, M, SM, R>), dim3((unsigned)nwg), dim3(GB), 0, \
                           st, src, S, snk, K, G, cell_start, rc4, bx4, mq, info->nmacro, order, \
                           hb, dist, skip_self, pcount, pq, PAR_QUEUE_CAP, perm ? 1 : 0);      \
    } while (0)
#define LEM_VGL2(A, SM, R, WP)                                                                \
    do {                                                                                      \
        if (sink_major)                                                                       \
            LEM_VGL(A, true, SM, R, WP);                                                      \
        else                                                                                  \
            LEM_VGL(A, false, SM, R, WP);                                                     \
    } while (0)
        if (semantics == LEM_SEM_NATIVE) {
            if (rev)
                LEM_VGL2(true, 1, true, LDS_WG_PER_CU);
            else
                LEM_VGL2(true, 1, false, LDS_WG_PER_CU);
        } else if (any) {
            if (rev && w7)
                LEM_VGL2(true, 0, true, 7);
            else if (rev)
                LEM_VGL2(true, 0, true, LDS_WG_PER_CU);
            else if (w7)
                LEM_VGL2(true, 0, false, 7);
            else
                LEM_VGL2(true, 0, false, LDS_WG_PER_CU);
        } else {
            LEM_VGL2(false, 0, false, LDS_WG_PER_CU);
        }
#undef LEM_VGL2
#undef LEM_VGL
        int rc = check_launch("k_visibility_grid_lds / k_visibility_grid_par");
        if (rc || !perm)
            return rc;
        const int64_t nub = std::max<int64_t>(1, std::min<int64_t>(NO, (int64_t)cus * 8));
        hipLaunchKernelGGL(k_unpermute_rows, dim3((unsigned)nub), dim3(GB),
                           (unsigned)(Wn * sizeof(unsigned long long)), st, hb, NO, Wn, NI, L.v0);
        return check_launch("k_unpermute_rows");
    }
    LEM_REQUIRE(NO * nbl <= 2147483647LL, "lem_visibility_grid: S*K too large for one launch");
    dim3 grid((unsigned)(NO * nbl));
#define LEM_VG(A, M, SM, R)                                                                   \
    hipLaunchKernelGGL((k_visibility_grid<A, M, SM, R>), grid, dim3(GB), 0, st, src, S, snk, K,   \
                       G, cell_start, rc4, bx4, order, hb, dist, nbl, skip_self)
#define LEM_VG2(A, SM, R)                                                                     \
    do {                                                                                      \
        if (sink_major)                                                                       \
            LEM_VG(A, true, SM, R);                                                           \
        else                                                                                  \
            LEM_VG(A, false, SM, R);                                                          \
    } while (0)
    if (semantics == LEM_SEM_NATIVE) {
        if (rev)
            LEM_VG2(true, 1, true);
        else
            LEM_VG2(true, 1, false);
    } else if (any) {
        if (rev)
            LEM_VG2(true, 0, true);
        else
            LEM_VG2(true, 0, false);
    } else {
        LEM_VG2(false, 0, false);
    }
#undef LEM_VG2
#undef LEM_VG
    return check_launch("k_visibility_grid");
}

extern "C" int lem_visibility_grid_stats(const float *src, int64_t S, const float *snk, int64_t K,
                                         const lem_grid_info *info, const int32_t *cell_start,
                                         const float *recs, const float *cell_box, int flags,
                                         uint64_t *stats, void *stream)
{
    clear_error();
    LEM_REQUIRE(S > 0 && K > 0 && src && snk && info && cell_start && recs && cell_box && stats,
                "lem_visibility_grid_stats: bad argument");
    const int64_t nbl = ceil_div(S, GB);
    LEM_REQUIRE(K * nbl <= 2147483647LL, "lem_visibility_grid_stats: S*K too large");
    GridDev G = to_dev(*info);
    hipStream_t st = as_stream(stream);
    unsigned long long *sp = reinterpret_cast<unsigned long long *>(stats);
    const float4 *rc4 = reinterpret_cast<const float4 *>(recs);
    const float4 *bx4 = reinterpret_cast<const float4 *>(cell_box);
    const uint2 *mq = reinterpret_cast<const uint2 *>(cell_box + 8 * (info->ncells + info->nmacro));
    const int nm = info->nmacro;
    if (flags & LEM_VIS_ANY_HIT)
        hipLaunchKernelGGL((k_grid_stats<true>), dim3((unsigned)(K * nbl)), dim3(GB), 0, st, src,
                           S, snk, K, G, cell_start, rc4, bx4, mq, nm, sp, nbl);
    else
        hipLaunchKernelGGL((k_grid_stats<false>), dim3((unsigned)(K * nbl)), dim3(GB), 0, st, src,
                           S, snk, K, G, cell_start, rc4, bx4, mq, nm, sp, nbl);
    return check_launch("k_grid_stats");
}

extern "C" int64_t lem_field_los_grid_workspace(int64_t S, int64_t K, int flags)
{
    if (S < 0 || K < 0)
        return -1;
    if (S == 0 || K == 0)
        return 0;
    const int64_t part = ((2 * los_chunk_count(S, K) * 6 * K * (int64_t)sizeof(float) + 255) / 256) *
                         256;
    if (flags & LEM_FIELD_INPUT_ORDER)
        return part;
    if (K > 2147483647LL)
        return -1;
    const int64_t cub = vis_sort_bytes(K);
    if (cub < 0)
        return -1;
    return part + vis_layout(K, (size_t)cub, nullptr, nullptr);
}

extern "C" int lem_field_los_grid(const float *src, const float *src_w, int64_t S,
                                  const float *snk, const float *snk_n, const float *snk_w,
                                  int64_t K, const lem_grid_info *info, const int32_t *cell_start,
                                  const float *recs, const float *cell_box, double alpha,
                                  double beta, int flags,
                                  float *out, void *workspace, int64_t workspace_bytes,
                                  void *stream)
{
    clear_error();
    LEM_REQUIRE(S > 0 && K > 0, "lem_field_los_grid: empty source or sink set");
    LEM_REQUIRE(src && src_w && snk && snk_n && info && cell_start && recs && cell_box && out,
                "lem_field_los_grid: null pointer");
    LEM_REQUIRE(!(flags & LEM_FIELD_ELEMENTS), "lem_field_los_grid: source sum only");
    const int64_t NC = los_chunk_count(S, K);
    const int64_t need = lem_field_los_grid_workspace(S, K, flags);
    if (need < 0) {
        set_error("lem_field_los_grid: workspace size query failed");
        return E_HIP;
    }
    if (!workspace || workspace_bytes < need) {
        set_error("lem_field_los_grid: workspace smaller than lem_field_los_grid_workspace()");
        return E_WORKSPACE;
    }
    const int64_t CH = ceil_div(S, NC);
    const int64_t nbk = ceil_div(K, FB);
    LEM_REQUIRE(NC * nbk <= 2147483647LL, "lem_field_los_grid: grid too large");
    float *ws_sum = reinterpret_cast<float *>(workspace);
    float *ws_cmp = ws_sum + NC * 6 * K;
    hipStream_t st = as_stream(stream);
    const int32_t *order = nullptr;
    if (!(flags & LEM_FIELD_INPUT_ORDER)) {
        const int64_t part = ((2 * NC * 6 * K * (int64_t)sizeof(float) + 255) / 256) * 256;
        VisWs L;
        vis_layout(K, (size_t)vis_sort_bytes(K), reinterpret_cast<char *>(workspace) + part, &L);
        int rc = morton_order(snk, K, L, st);
        if (rc)
            return rc;
        order = L.order;
    }
    if (alpha != 0.0)
        hipLaunchKernelGGL(k_field_los_grid_partial<true>, dim3((unsigned)(NC * nbk)), dim3(FB), 0, st, src,
                       src_w, S, snk, snk_n, K, to_dev(*info), cell_start,
                       reinterpret_cast<const float4 *>(recs),
                       reinterpret_cast<const float4 *>(cell_box), order, (float)alpha, beta, CH,
                       nbk, ws_sum, ws_cmp);
    else
        hipLaunchKernelGGL(k_field_los_grid_partial<false>, dim3((unsigned)(NC * nbk)), dim3(FB), 0, st, src,
                       src_w, S, snk, snk_n, K, to_dev(*info), cell_start,
                       reinterpret_cast<const float4 *>(recs),
                       reinterpret_cast<const float4 *>(cell_box), order, (float)alpha, beta, CH,
                       nbk, ws_sum, ws_cmp);
    int rc = check_launch("k_field_los_grid_partial");
    if (rc)
        return rc;
    return launch_los_finish(ws_sum, ws_cmp, NC, K, snk_n, snk_w,
                             (flags & LEM_FIELD_SINK_TANGENT) ? 1 : 0, out, st);
}
