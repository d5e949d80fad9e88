// lem_field_dev.h -- per-pair field helpers shared by the field kernels
// (masked by a precomputed bitmask in lem_fields.hip, fused with the grid
// traversal in lem_grid.hip).
#pragma once

#include "lem_device.h"

namespace lem {

constexpr int FB = 256;       // sinks per block
constexpr int FS_TILE = 256;  // sources per LDS tile

struct SrcLds {
    double x, y, z;
    float w[6];
};

LEM_DEV C3 load_w(const float *__restrict__ w, int64_t i)
{
    const float2 *p = reinterpret_cast<const float2 *>(w) + 3 * i;
    float2 a = p[0], b = p[1], c = p[2];
    C3 o;
    o.xr = a.x; o.xi = a.y; o.yr = b.x; o.yi = b.y; o.zr = c.x; o.zi = c.y;
    return o;
}

LEM_DEV C3 cmul_comp(const C3 &a, const C3 &b)
{
#pragma clang fp contract(fast)
    C3 o;
    o.xr = a.xr * b.xr - a.xi * b.xi; o.xi = a.xr * b.xi + a.xi * b.xr;
    o.yr = a.yr * b.yr - a.yi * b.yi; o.yi = a.yr * b.yi + a.yi * b.yr;
    o.zr = a.zr * b.zr - a.zi * b.zi; o.zi = a.zr * b.zi + a.zi * b.zr;
    return o;
}

LEM_DEV C3 scale(const C3 &a, float lr, float li)
{
#pragma clang fp contract(fast)
    C3 o;
    o.xr = a.xr * lr - a.xi * li; o.xi = a.xr * li + a.xi * lr;
    o.yr = a.yr * lr - a.yi * li; o.yi = a.yr * li + a.yi * lr;
    o.zr = a.zr * lr - a.zi * li; o.zi = a.zr * li + a.zi * lr;
    return o;
}

// Projector onto the plane normal to n (normalised here; the reference builds
// its u, v basis from normalised cross products, empropagation.py:1150-1156).
LEM_DEV C3 project_normal(const C3 &e, float nx, float ny, float nz)
{
    float inv = 1.0f / sqrtf(nx * nx + ny * ny + nz * nz);
    return project(e, nx * inv, ny * inv, nz * inv);
}

// One source->sink leg before the sink weight / tangent step: the projected
// source weight e = (I - d d^T) w_s and the scalar leg factor (lr, li); the
// contribution is (lr + j li) e.
template <bool LOSSY = true>
LEM_DEV void los_leg(double xs, double ys, double zs, const C3 &ws, double xk, double yk,
                     double zk, float nkx, float nky, float nkz, float alpha, double beta, C3 &e,
                     float &lr, float &li)
{
#pragma clang fp contract(fast)
    double dx = xk - xs, dy = yk - ys, dz = zk - zs;
    Phase p = phase_of(dx, dy, dz, beta);
    float ux = (float)dx * p.inv_r, uy = (float)dy * p.inv_r, uz = (float)dz * p.inv_r;
    float cosang = ux * nkx + uy * nky + uz * nkz;
    leg_factor<LOSSY>(p, alpha, (float)beta, cosang, lr, li);
    e = project(ws, ux, uy, uz);
}

template <bool LOSSY = true>
LEM_DEV C3 los_term(double xs, double ys, double zs, const C3 &ws, double xk, double yk,
                    double zk, float nkx, float nky, float nkz, float alpha, double beta)
{
    C3 e;
    float lr, li;
    los_leg<LOSSY>(xs, ys, zs, ws, xk, yk, zk, nkx, nky, nkz, alpha, beta, e, lr, li);
    return scale(e, lr, li);
}

// g += (lr + j li) e, each product fused into the running partial (the same
// operations in every kernel that sums legs, so their sums agree bit for bit).
LEM_DEV void add_leg(float g[6], const C3 &e, float lr, float li)
{
    g[0] = __builtin_fmaf(e.xr, lr, g[0]); g[0] = __builtin_fmaf(-e.xi, li, g[0]);
    g[1] = __builtin_fmaf(e.xr, li, g[1]); g[1] = __builtin_fmaf(e.xi, lr, g[1]);
    g[2] = __builtin_fmaf(e.yr, lr, g[2]); g[2] = __builtin_fmaf(-e.yi, li, g[2]);
    g[3] = __builtin_fmaf(e.yr, li, g[3]); g[3] = __builtin_fmaf(e.yi, lr, g[3]);
    g[4] = __builtin_fmaf(e.zr, lr, g[4]); g[4] = __builtin_fmaf(-e.zi, li, g[4]);
    g[5] = __builtin_fmaf(e.zr, li, g[5]); g[5] = __builtin_fmaf(e.zi, lr, g[5]);
}

LEM_DEV bool blocked(const unsigned long long *__restrict__ hit_bits, int64_t W64, int64_t s,
                     int64_t k)
{
    if (!hit_bits)
        return false;
    unsigned long long w = hit_bits[s * W64 + (k >> 6)];
    return (w >> (k & 63)) & 1ull;
}

// Mask bit of ray s -> k in either layout: source-major (S x W64 words over
// sinks) or sink-major (K x W64 words over sources).
LEM_DEV bool blocked_any(const unsigned long long *__restrict__ hit_bits, int64_t W64,
                         bool sink_major, int64_t s, int64_t k)
{
    return sink_major ? blocked(hit_bits, W64, k, s) : blocked(hit_bits, W64, s, k);
}

LEM_DEV void stage_sources(SrcLds *s_src, const float *__restrict__ src,
                           const float *__restrict__ src_w, int64_t base, int64_t s_end)
{
    int64_t j = base + threadIdx.x;
    SrcLds L;
    if (j < s_end) {
        L.x = (double)src[3 * j];
        L.y = (double)src[3 * j + 1];
        L.z = (double)src[3 * j + 2];
        C3 w = load_w(src_w, j);
        L.w[0] = w.xr; L.w[1] = w.xi; L.w[2] = w.yr; L.w[3] = w.yi; L.w[4] = w.zr; L.w[5] = w.zi;
    } else {
        L.x = L.y = L.z = 0.0;
        for (int i = 0; i < 6; ++i)
            L.w[i] = 0.0f;
    }
    s_src[threadIdx.x] = L;
}

}  // namespace lem
