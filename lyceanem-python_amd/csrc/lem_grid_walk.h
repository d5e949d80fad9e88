// lem_grid_walk.h -- the uniform grid's layout, the per-item arithmetic of its
// build and the two-level DDA walk (lem_grid.hip's traversal kernels call
// dda_cast).  Everything here is plain per-element code, so the same source
// also compiles as host C++ (LEM_HOST_BUILD): a CPU build of the walk and of
// the build's index arithmetic checks them against the oracle's brute force,
// also under AddressSanitizer / UndefinedBehaviorSanitizer.
#pragma once

#include "lem_device.h"
#include "../../include/lem.h"

#include <math.h>
#include <stdint.h>

#include <algorithm>

namespace lem {

// Macro blocks hold 2^mshift cells per axis: the smallest shift from
// LEM_GRID_MSHIFT up for which the quantised macro table (8 bytes a block)
// fits the LDS budget of the persistent traversal.
#ifndef LEM_GRID_MSHIFT
#define LEM_GRID_MSHIFT 2
#endif
constexpr int MACRO_LDS_MAX = 4352;  // blocks (6 B each, 25.5 KB of LDS: 6 workgroups per CU)

// The quantised macro table as the walk reads it from LDS: per block the
// 32-bit word (lo.x, lo.y, lo.z, hi.x) and the 16-bit word (hi.y, hi.z) of
// the 8-byte table entry, kept in two arrays so a block takes 6 bytes.
// Beside it, the per-axis grid constants the walk's steps pick by axis (one
// 16-byte LDS read instead of selecting among three scalar registers, which a
// vector instruction can read only one at a time).
struct MacroTab {
    const uint32_t *a;
    const uint16_t *b;
    const float4 *axis;  // per axis: lo, cell, cells (int bits), macro blocks (int bits)
    LEM_DEV uint2 operator[](int m) const { return make_uint2(a[m], (uint32_t)b[m]); }
};

struct GridDev {
    float lox, loy, loz;
    float cx, cy, cz;     // cell size
    float icx, icy, icz;  // 1 / cell size
    int nx, ny, nz;
    float eps;
    int mshift;           // macro blocks of 2^mshift cells per axis
    int mx, my, mz;       // macro blocks per axis
    float qlx, qly, qlz;  // lo - eps: origin of the quantised macro boxes
    float qx, qy, qz;     // (macro-block extent + 2 eps) / 255: their step
    int64_t ncells;       // macro boxes follow the cell boxes
};

inline GridDev to_dev(const lem_grid_info &g)
{
    GridDev d;
    d.lox = g.lo[0]; d.loy = g.lo[1]; d.loz = g.lo[2];
    d.cx = g.cell[0]; d.cy = g.cell[1]; d.cz = g.cell[2];
    d.icx = 1.0f / g.cell[0]; d.icy = 1.0f / g.cell[1]; d.icz = 1.0f / g.cell[2];
    d.nx = g.n[0]; d.ny = g.n[1]; d.nz = g.n[2];
    d.eps = g.eps;
    d.mshift = g.mshift;
    const int B = 1 << g.mshift;
    d.mx = (g.n[0] + B - 1) / B;
    d.my = (g.n[1] + B - 1) / B;
    d.mz = (g.n[2] + B - 1) / B;
    // a quantised box spans its block widened by eps on each side, so a face
    // clamped to the block still decodes at least eps outside the block's own
    // face: the walk tests hit points up to a rounding error beyond the block
    // it is in, and rays along a block face (a zero direction component) are
    // tested against the box with no margin
    d.qlx = d.lox - d.eps;
    d.qly = d.loy - d.eps;
    d.qlz = d.loz - d.eps;
    d.qx = (d.cx * (float)B + 2.0f * d.eps) / 255.0f;
    d.qy = (d.cy * (float)B + 2.0f * d.eps) / 255.0f;
    d.qz = (d.cz * (float)B + 2.0f * d.eps) / 255.0f;
    d.ncells = g.ncells;
    return d;
}

LEM_DEV int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Entry a of MacroTab::axis.
LEM_DEV float4 axis_entry(const GridDev &G, int a)
{
    return a == 0 ? make_float4(G.lox, G.cx, __builtin_bit_cast(float, G.nx),
                                __builtin_bit_cast(float, G.mx))
         : a == 1 ? make_float4(G.loy, G.cy, __builtin_bit_cast(float, G.ny),
                                __builtin_bit_cast(float, G.my))
                  : make_float4(G.loz, G.cz, __builtin_bit_cast(float, G.nz),
                                __builtin_bit_cast(float, G.mz));
}

// Grid constants of the axis a step crosses (ax: x, else ay: y, else z).
struct AxisK {
    float lo, c;
    int n, nm;
};

template <bool LDSM>
LEM_DEV AxisK axis_k(const GridDev &G, const MacroTab &smq, bool ax, bool ay)
{
    if constexpr (LDSM) {
        const float4 k = smq.axis[ax ? 0 : (ay ? 1 : 2)];
        return AxisK{k.x, k.y, __builtin_bit_cast(int, k.z), __builtin_bit_cast(int, k.w)};
    } else {
        return AxisK{ax ? G.lox : (ay ? G.loy : G.loz), ax ? G.cx : (ay ? G.cy : G.cz),
                     ax ? G.nx : (ay ? G.ny : G.nz), ax ? G.mx : (ay ? G.my : G.mz)};
    }
}

LEM_DEV void tri_cell_range(const GridDev &G, const float *t9, int lo[3], int hi[3])
{
    float mn[3], mx[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float v0 = t9[a], v1 = t9[3 + a], v2 = t9[6 + a];
        mn[a] = fminf(v0, fminf(v1, v2)) - G.eps;
        mx[a] = fmaxf(v0, fmaxf(v1, v2)) + G.eps;
    }
    lo[0] = clampi((int)floorf((mn[0] - G.lox) * G.icx), 0, G.nx - 1);
    lo[1] = clampi((int)floorf((mn[1] - G.loy) * G.icy), 0, G.ny - 1);
    lo[2] = clampi((int)floorf((mn[2] - G.loz) * G.icz), 0, G.nz - 1);
    hi[0] = clampi((int)floorf((mx[0] - G.lox) * G.icx), 0, G.nx - 1);
    hi[1] = clampi((int)floorf((mx[1] - G.loy) * G.icy), 0, G.ny - 1);
    hi[2] = clampi((int)floorf((mx[2] - G.loz) * G.icz), 0, G.nz - 1);
}

LEM_DEV float macro_origin(float lo, float c, int j, int ms) { return lo + c * (float)(j << ms); }

// ---------------------------------------------------------------------------
// DDA traversal
// ---------------------------------------------------------------------------

// Triangle id of record i (kept in the record's padding).
LEM_DEV int32_t rec_id(const float4 *__restrict__ recs, int32_t i)
{
    return __builtin_bit_cast(int32_t, recs[3 * (int64_t)i + 2].y);
}

LEM_DEV Tri load_rec(const float4 *__restrict__ recs, int32_t i)
{
    float4 a = recs[3 * (int64_t)i], b = recs[3 * (int64_t)i + 1], c = recs[3 * (int64_t)i + 2];
    Tri t;
    t.v0x = a.x; t.v0y = a.y; t.v0z = a.z; t.e1x = a.w;
    t.e1y = b.x; t.e1z = b.y; t.e2x = b.z; t.e2y = b.w;
    t.e2z = c.x;
    return t;
}

// One slab of the segment-box test, without branches: the t interval in
// which the ray is between lo and hi on this axis narrows [tn, tf].  A ray
// parallel to the axis (par; its inverse direction is infinite) is inside the
// slab for every t or for none.
LEM_DEV void slab(bool par, float o, float lo, float hi, float id, float &tn, float &tf)
{
    const float INF = __builtin_huge_valf();
    const float a = (lo - o) * id, b = (hi - o) * id;
    const bool inside = o >= lo && o <= hi;
    const float n = par ? (inside ? -INF : INF) : fminf(a, b);
    const float f = par ? (inside ? INF : -INF) : fmaxf(a, b);
    tn = fmaxf(tn, n);
    tf = fminf(tf, f);
}

// True when the segment t in [t_lo, t_hi] of the ray meets the box (slab test
// with a margin of m in t).  PAR = false: no direction component is zero.
template <bool PAR = true>
LEM_DEV bool seg_box(const Ray &r, float idx, float idy, float idz, float4 lo, float4 hi,
                     float t_lo, float t_hi, float m)
{
    float tn = t_lo, tf = t_hi;
    slab(PAR && r.dx == 0.0f, r.ox, lo.x, hi.x, idx, tn, tf);
    slab(PAR && r.dy == 0.0f, r.oy, lo.y, hi.y, idy, tn, tf);
    slab(PAR && r.dz == 0.0f, r.oz, lo.z, hi.z, idz, tn, tf);
    return tn <= tf + m;
}

// Distance along the ray to the far boundary of cell i on one axis (s = step
// direction, 0 when the ray is parallel to the axis), from the ray origin.
LEM_DEV float tnext(int i, int s, float lo, float c, float o, float id)
{
    if (s == 0)
        return __builtin_huge_valf();
    float bnd = lo + c * (float)(i + (s > 0 ? 1 : 0));
    return (bnd - o) * id;
}

// Start of a cell walk: the segment [0, len] clipped to the grid box
// ([t0, t1]), the cell holding the clipped start, step directions and the
// distances to the first boundary on each axis.
struct Walk {
    float idx, idy, idz;
    float t1;
    int ix, iy, iz, sx, sy, sz;
    float tx, ty, tz;
};

LEM_DEV bool walk_begin(const Ray &r, const GridDev &G, Walk &w)
{
    // a zero-length segment (source == sink: NaN direction) crosses no cell;
    // the predicate rejects every triangle for it too
    if (!(r.len > 0.0f))
        return false;
    float hx = G.lox + G.cx * G.nx, hy = G.loy + G.cy * G.ny, hz = G.loz + G.cz * G.nz;
    w.idx = 1.0f / r.dx;
    w.idy = 1.0f / r.dy;
    w.idz = 1.0f / r.dz;
    float t0 = 0.0f, t1 = r.len;
    float a = (G.lox - r.ox) * w.idx, b = (hx - r.ox) * w.idx;
    if (r.dx == 0.0f) {
        if (r.ox < G.lox || r.ox > hx) return false;
    } else {
        t0 = fmaxf(t0, fminf(a, b));
        t1 = fminf(t1, fmaxf(a, b));
    }
    a = (G.loy - r.oy) * w.idy; b = (hy - r.oy) * w.idy;
    if (r.dy == 0.0f) {
        if (r.oy < G.loy || r.oy > hy) return false;
    } else {
        t0 = fmaxf(t0, fminf(a, b));
        t1 = fminf(t1, fmaxf(a, b));
    }
    a = (G.loz - r.oz) * w.idz; b = (hz - r.oz) * w.idz;
    if (r.dz == 0.0f) {
        if (r.oz < G.loz || r.oz > hz) return false;
    } else {
        t0 = fmaxf(t0, fminf(a, b));
        t1 = fminf(t1, fmaxf(a, b));
    }
    if (!(t0 <= t1 + G.eps))
        return false;
    w.t1 = t1;
    float pm = t0 > 0.0f ? t0 : 0.0f;
    w.ix = clampi((int)floorf((r.ox + r.dx * pm - G.lox) * G.icx), 0, G.nx - 1);
    w.iy = clampi((int)floorf((r.oy + r.dy * pm - G.loy) * G.icy), 0, G.ny - 1);
    w.iz = clampi((int)floorf((r.oz + r.dz * pm - G.loz) * G.icz), 0, G.nz - 1);
    w.sx = r.dx > 0.0f ? 1 : (r.dx < 0.0f ? -1 : 0);
    w.sy = r.dy > 0.0f ? 1 : (r.dy < 0.0f ? -1 : 0);
    w.sz = r.dz > 0.0f ? 1 : (r.dz < 0.0f ? -1 : 0);
    w.tx = tnext(w.ix, w.sx, G.lox, G.cx, r.ox, w.idx);
    w.ty = tnext(w.iy, w.sy, G.loy, G.cy, r.oy, w.idy);
    w.tz = tnext(w.iz, w.sz, G.loz, G.cz, r.oz, w.idz);
    return true;
}

// kernel1Dv3's distance (rayfunctions.py:212-244) is order dependent: it
// tests triangles in index order and accepts a hit only below the current
// distance minus 2e-7, storing the distance as float32.  Given the exact
// closest hit t* (float64), only hits within a short chain of such windows
// above t* can change the outcome; hits farther than delta = 8 (2e-7 + ulp(t*))
// never reject one of them.  This replays the window from dist = len over
// every hit in [t*, t* + delta], in triangle-id order, taken from the cells
// the piece [t* - delta, t* + delta] of the ray passes through (each cell
// lists its triangles by ascending id; triangle boxes are inflated by eps,
// so each such hit's triangle is in one of these cells).
LEM_DEV float replay_window(const Ray &r, double tstar, const GridDev &G,
                            const int32_t *__restrict__ cell_start,
                            const float4 *__restrict__ recs)
{
    const float fs = (float)tstar;
    const float ulp = nextafterf(fs, __builtin_huge_valf()) - fs;
    const double delta = 8.0 * (2e-7 + (double)ulp);
    const double ta = tstar - delta > 0.0 ? tstar - delta : 0.0, tb = tstar + delta;
    const float o[3] = {r.ox, r.oy, r.oz}, d[3] = {r.dx, r.dy, r.dz};
    const float glo[3] = {G.lox, G.loy, G.loz}, gic[3] = {G.icx, G.icy, G.icz};
    const int gn[3] = {G.nx, G.ny, G.nz};
    int lo[3], hi[3];
    int ncell = 1;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float pa = o[a] + d[a] * (float)ta, pb = o[a] + d[a] * (float)tb;
        const int ca = clampi((int)floorf((pa - glo[a]) * gic[a]), 0, gn[a] - 1);
        const int cb = clampi((int)floorf((pb - glo[a]) * gic[a]), 0, gn[a] - 1);
        lo[a] = min(ca, cb);
        hi[a] = max(ca, cb);
        ncell *= hi[a] - lo[a] + 1;
    }
    if (ncell > 64)
        return fs;  // a piece across more than 64 cells: keep the closest hit
    float D = r.len;
    int last = -1;
    for (int it = 0; it < 64; ++it) {
        // next hit of the piece in id order
        int best = 0x7fffffff;
        double bt = 0.0;
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    const int c = (z * G.ny + y) * G.nx + x;
                    for (int32_t j = cell_start[c]; j < cell_start[c + 1]; ++j) {
                        const int id = rec_id(recs, j);
                        if (id <= last)
                            continue;
                        if (id >= best)
                            break;
                        double tt;
                        if (hit_numba(r, r.len, load_rec(recs, j), tt) && tt <= tb) {
                            best = id;
                            bt = tt;
                            break;
                        }
                    }
                }
        if (best == 0x7fffffff)
            break;
        if (bt < (double)D - 2e-7)
            D = (float)bt;
        last = best;
    }
    return D;
}

// Casts one charged ray through the grid.  Returns the blocked flag and
// leaves kernel1Dv3's `dist` in `dist` (closest-hit mode).
//
// Two-level walk.  Macro blocks (2^mshift cells per axis) are stepped with
// their own boundary distances T*, taken from the same tnext formula at the
// block's last cell, so a block face is the identical float32 value at both
// levels.  A block whose content box the segment misses is left in one step;
// inside a live block the cells are walked one by one and each cell's
// triangles are tested when the segment meets the cell's box.  On entering a
// live block after skipping, the entry cell on the two axes not crossed comes
// from the entry point (clamped into the block; a cell off by one at an edge
// or corner only shifts the walk within the eps inflation of the boxes).
// LDSM: macro boxes come from the quantised table `smq` (LDS) instead of the
// float boxes in global memory.
// rt (any-hit only): the ray whose predicate decides; the walk then covers
// the same segment with r (e.g. charged from the other end, to reach hits
// near the sink first) and every triangle is tested with *rt.  The mask of
// any-hit mode does not depend on the order of the tests.
template <bool ANY, int SEM = 0, bool LDSM = false, bool PAR = true>
LEM_DEV bool dda_cast(const Ray &r, float &dist, const GridDev &G,
                      const int32_t *__restrict__ cell_start, const float4 *__restrict__ recs,
                      const float4 *__restrict__ box, unsigned int *n_tests = nullptr,
                      unsigned int *n_cells = nullptr, unsigned int *n_macro = nullptr,
                      MacroTab smq = MacroTab{nullptr, nullptr, nullptr},
                      const Ray *rt = nullptr)
{
    const Ray &rq = rt ? *rt : r;  // the predicate's ray
    Walk w;
    if (!walk_begin(r, G, w))
        return false;
    const float idx = w.idx, idy = w.idy, idz = w.idz, t1 = w.t1;
    const int sx = w.sx, sy = w.sy, sz = w.sz;
    int ix = w.ix, iy = w.iy, iz = w.iz;
    float tx = w.tx, ty = w.ty, tz = w.tz;
    const float margin = 2.0f * G.eps;
    const int ms = G.mshift, B = 1 << ms;
    const float4 *__restrict__ mbox = box + 2 * G.ncells;  // macro-block boxes
    // Quantised boxes, rays with no zero direction component: the slab bound
    // of face q of block j on an axis is t = A + C j + S q, with
    // A = (lo - eps - o) / d, C = cell * 2^mshift / d, S = step / d (the same
    // quantities as the decoded box, reassociated; the 2 eps margin in t
    // covers the rounding difference).
    const float Axx = (G.qlx - r.ox) * idx, Ayy = (G.qly - r.oy) * idy,
                Azz = (G.qlz - r.oz) * idz;
    const float Cxx = G.cx * (float)B * idx, Cyy = G.cy * (float)B * idy,
                Czz = G.cz * (float)B * idz;
    const float Sxx = G.qx * idx, Syy = G.qy * idy, Szz = G.qz * idz;
    // last cell of block j along an axis in the direction of travel
    auto last = [&](int j, int s, int n) { return s > 0 ? min((j << ms) + B - 1, n - 1) : (j << ms); };
    int jx = ix >> ms, jy = iy >> ms, jz = iz >> ms;
    float Tx = tnext(last(jx, sx, G.nx), sx, G.lox, G.cx, r.ox, idx);
    float Ty = tnext(last(jy, sy, G.ny), sy, G.loy, G.cy, r.oy, idy);
    float Tz = tnext(last(jz, sz, G.nz), sz, G.loz, G.cz, r.oz, idz);
    bool hit = false;
    double tmin = (double)r.len;  // closest hit so far (closest-hit mode)
    double t_next = __builtin_huge_val();  // closest hit of another triangle
    int32_t id_min = -1;
    bool fine = true;   // (ix, iy, iz) is the walk's cell in block (jx, jy, jz)
    int entry_axis = 0;
    float t_entry = 0.0f;
    // Every iteration moves a block index one step in the direction of travel
    // or walks cells inside one block (at most 3 * 2^mshift steps); block indices
    // never move back, so the walk ends; the caps only guard the impossible.
    for (int guard = G.mx + G.my + G.mz + 4; guard > 0; --guard) {
        const int m = (jz * G.my + jy) * G.mx + jx;
        bool live;
        if (LDSM && !PAR) {
            const uint2 qv = smq[m];
            // fused multiply-adds: the eps-widened boxes and the 2 eps margin in t
            // absorb the difference in rounding
            const float bx = __builtin_fmaf(Cxx, (float)jx, Axx),
                        by = __builtin_fmaf(Cyy, (float)jy, Ayy),
                        bz = __builtin_fmaf(Czz, (float)jz, Azz);
            const float x0 = __builtin_fmaf(Sxx, (float)(qv.x & 255u), bx),
                        x1 = __builtin_fmaf(Sxx, (float)(qv.x >> 24), bx);
            const float y0 = __builtin_fmaf(Syy, (float)((qv.x >> 8) & 255u), by),
                        y1 = __builtin_fmaf(Syy, (float)(qv.y & 255u), by);
            const float z0 = __builtin_fmaf(Szz, (float)((qv.x >> 16) & 255u), bz),
                        z1 = __builtin_fmaf(Szz, (float)((qv.y >> 8) & 255u), bz);
            const float tn = fmaxf(fmaxf(-margin, fminf(x0, x1)), fmaxf(fminf(y0, y1), fminf(z0, z1)));
            const float tf = fminf(fminf(dist + margin, fmaxf(x0, x1)),
                                   fminf(fmaxf(y0, y1), fmaxf(z0, z1)));
            live = ((qv.x & 255u) <= (qv.x >> 24)) & (tn <= tf + margin);
        } else {
            float4 mlo, mhi;
            if (LDSM) {
                const uint2 qv = smq[m];
                const float ox = macro_origin(G.qlx, G.cx, jx, ms),
                            oy = macro_origin(G.qly, G.cy, jy, ms),
                            oz = macro_origin(G.qlz, G.cz, jz, ms);
                mlo = make_float4(ox + (float)(qv.x & 255u) * G.qx,
                                  oy + (float)((qv.x >> 8) & 255u) * G.qy,
                                  oz + (float)((qv.x >> 16) & 255u) * G.qz, 0.0f);
                mhi = make_float4(ox + (float)(qv.x >> 24) * G.qx, oy + (float)(qv.y & 255u) * G.qy,
                                  oz + (float)((qv.y >> 8) & 255u) * G.qz, 0.0f);
            } else {
                // both halves of the box are loaded before either is tested
                mlo = mbox[2 * m];
                mhi = mbox[2 * m + 1];
            }
            live = (mlo.x <= mhi.x) &
                   seg_box<PAR>(r, idx, idy, idz, mlo, mhi, -margin, dist + margin, margin);
        }
        if (n_macro)
            *n_macro += 1;
        if (!live) {
            // leave the block across its nearest face; the axis is selected,
            // not branched on, so lanes crossing different axes stay together
            const bool ax = Tx <= Ty && Tx <= Tz;
            const bool ay = !ax && Ty <= Tz;
            const float tex = ax ? Tx : (ay ? Ty : Tz);
            const int sa = ax ? sx : (ay ? sy : sz);
            const int ja = (ax ? jx : (ay ? jy : jz)) + sa;
            const AxisK K = axis_k<LDSM>(G, smq, ax, ay);
            // the exits are combined into one test (one branch per step)
            if ((tex >= t1 + margin) | (hit & (dist + margin <= tex)) | (ja < 0) | (ja >= K.nm))
                break;
            const float Ta = tnext(last(ja, sa, K.n), sa, K.lo, K.c, ax ? r.ox : (ay ? r.oy : r.oz),
                                   ax ? idx : (ay ? idy : idz));
            jx = ax ? ja : jx;
            jy = ay ? ja : jy;
            jz = (!ax && !ay) ? ja : jz;
            Tx = ax ? Ta : Tx;
            Ty = ay ? Ta : Ty;
            Tz = (!ax && !ay) ? Ta : Tz;
            entry_axis = ax ? 0 : (ay ? 1 : 2);
            t_entry = tex;
            fine = false;
            continue;
        }
        if (!fine) {
            // first cell of the block: the crossed axis enters at the block's
            // first cell in the direction of travel, the others at the entry point
            const int bx = jx << ms, by = jy << ms, bz = jz << ms;
            const int ex = min(bx + B - 1, G.nx - 1), ey = min(by + B - 1, G.ny - 1),
                      ez = min(bz + B - 1, G.nz - 1);
            const float px = r.ox + r.dx * t_entry, py = r.oy + r.dy * t_entry,
                        pz = r.oz + r.dz * t_entry;
            ix = clampi((int)floorf((px - G.lox) * G.icx), bx, ex);
            iy = clampi((int)floorf((py - G.loy) * G.icy), by, ey);
            iz = clampi((int)floorf((pz - G.loz) * G.icz), bz, ez);
            if (entry_axis == 0)
                ix = sx > 0 ? bx : ex;
            else if (entry_axis == 1)
                iy = sy > 0 ? by : ey;
            else
                iz = sz > 0 ? bz : ez;
            tx = tnext(ix, sx, G.lox, G.cx, r.ox, idx);
            ty = tnext(iy, sy, G.loy, G.cy, r.oy, idy);
            tz = tnext(iz, sz, G.loz, G.cz, r.oz, idz);
            fine = true;
        }
        bool done = false;
        for (int g = 3 * B + 3; g > 0; --g) {
            int32_t b, e;
            {
                const int c = (iz * G.ny + iy) * G.nx + ix;
                // box and record range of the cell in one 32-byte read
                const float4 clo = box[2 * c], chi = box[2 * c + 1];
                b = __builtin_bit_cast(int32_t, clo.w);
                e = __builtin_bit_cast(int32_t, chi.w);
                if (n_cells)
                    *n_cells += 1;
                // tested unconditionally (an empty cell's range stays empty
                // whatever the test says), so the range and the box arrive in
                // one round trip instead of the box being fetched after the
                // range is known
                e = seg_box<PAR>(r, idx, idy, idz, clo, chi, -margin, dist + margin, margin) ? e
                                                                                              : b;
            }
            if (n_tests)
                *n_tests += (unsigned int)(e - b);
            for (int32_t j = b; j < e; ++j) {
                Tri t = load_rec(recs, j);
                if (SEM == 1) {
                    if (hit_native(rq, t))
                        return true;
                    continue;
                }
                double tt;
                if (ANY) {
                    if (hit_numba(rq, rt ? rq.len : dist, t, tt)) {
                        dist = (float)tt;
                        return true;
                    }
                } else if (hit_numba(r, r.len, t, tt)) {
                    // exact closest hit and the closest hit of any other
                    // triangle (a triangle listed in several cells counts once)
                    const int32_t id = rec_id(recs, j);
                    if (id != id_min) {
                        if (tt < tmin) {
                            if (hit)
                                t_next = fmin(t_next, tmin);
                            tmin = tt;
                            id_min = id;
                            dist = (float)tt;
                            hit = true;
                        } else if (tt < t_next) {
                            t_next = tt;
                        }
                    }
                }
            }
            const float texit = fminf(tx, fminf(ty, tz));
            // next cell across the nearest face (axis selected, as above); the
            // exits are combined into one test
            const bool ax = tx <= ty && tx <= tz;
            const bool ay = !ax && ty <= tz;
            const bool az = !ax && !ay;
            const int sa = ax ? sx : (ay ? sy : sz);
            const int ia = (ax ? ix : (ay ? iy : iz)) + sa;
            const AxisK K = axis_k<LDSM>(G, smq, ax, ay);
            const int na = K.n;
            if ((hit & (dist + margin <= texit)) | (texit >= t1 + margin) | (ia < 0) | (ia >= na)) {
                done = true;
                break;
            }
            const float loa = K.lo;
            const float ca = K.c;
            const float oa = ax ? r.ox : (ay ? r.oy : r.oz);
            const float ida = ax ? idx : (ay ? idy : idz);
            const float ta = tnext(ia, sa, loa, ca, oa, ida);
            ix = ax ? ia : ix;
            iy = ay ? ia : iy;
            iz = az ? ia : iz;
            tx = ax ? ta : tx;
            ty = ay ? ta : ty;
            tz = az ? ta : tz;
            const int ja = ia >> ms;
            if (ja != (ax ? jx : (ay ? jy : jz))) {
                // into the next block: its far boundary on this axis
                const float Ta = tnext(last(ja, sa, na), sa, loa, ca, oa, ida);
                jx = ax ? ja : jx;
                jy = ay ? ja : jy;
                jz = az ? ja : jz;
                Tx = ax ? Ta : Tx;
                Ty = ay ? Ta : Ty;
                Tz = az ? Ta : Tz;
                break;
            }
        }
        if (done)
            break;
    }
    if (!ANY && SEM == 0 && hit) {
        // another triangle's hit close enough to matter: replay kernel1Dv3's
        // window; alone, the closest hit is the answer
        const float fs = (float)tmin;
        const double delta = 8.0 * (2e-7 + (double)(nextafterf(fs, __builtin_huge_valf()) - fs));
        if (t_next <= tmin + delta)
            dist = replay_window(r, tmin, G, cell_start, recs);
    }
    return hit;
}

// Work of a conservative front-to-back DDA: the triangle references of every
// cell the segment crosses from its start up to the cell holding the first hit
// at distance t_stop (no box or macro-block pruning).  Measurement only.
LEM_DEV unsigned int conservative_refs(const Ray &r, float t_stop, const GridDev &G,
                                       const int32_t *__restrict__ cell_start)
{
    Walk w;
    if (!walk_begin(r, G, w))
        return 0;
    const float t_end = fminf(w.t1, t_stop);
    unsigned int n = 0;
    for (int guard = G.nx + G.ny + G.nz + 4; guard > 0; --guard) {
        const int c = (w.iz * G.ny + w.iy) * G.nx + w.ix;
        n += (unsigned int)(cell_start[c + 1] - cell_start[c]);
        if (w.tx <= w.ty && w.tx <= w.tz) {
            if (w.tx >= t_end) break;
            w.ix += w.sx;
            if (w.ix < 0 || w.ix >= G.nx) break;
            w.tx = tnext(w.ix, w.sx, G.lox, G.cx, r.ox, w.idx);
        } else if (w.ty <= w.tz) {
            if (w.ty >= t_end) break;
            w.iy += w.sy;
            if (w.iy < 0 || w.iy >= G.ny) break;
            w.ty = tnext(w.iy, w.sy, G.loy, G.cy, r.oy, w.idy);
        } else {
            if (w.tz >= t_end) break;
            w.iz += w.sz;
            if (w.iz < 0 || w.iz >= G.nz) break;
            w.tz = tnext(w.iz, w.sz, G.loz, G.cz, r.oz, w.idz);
        }
    }
    return n;
}

// Cell size from the scene extent and a target cell count: cubic cells over
// the axes thicker than a cell, a single layer along thinner axes.
inline void choose_resolution(const float ext[3], double target_cells, int64_t cap, int n[3])
{
    bool thin[3] = {false, false, false};
    double h = 0.0;
    for (int it = 0; it < 3; ++it) {
        double vol = 1.0;
        int dims = 0;
        for (int a = 0; a < 3; ++a)
            if (!thin[a]) {
                vol *= std::max((double)ext[a], 1e-30);
                ++dims;
            }
        if (dims == 0) {
            h = 1.0;
            break;
        }
        h = pow(vol / std::max(target_cells, 1.0), 1.0 / dims);
        bool changed = false;
        for (int a = 0; a < 3; ++a)
            if (!thin[a] && ext[a] < h) {
                thin[a] = true;
                changed = true;
            }
        if (!changed)
            break;
    }
    for (;;) {
        int64_t tot = 1;
        for (int a = 0; a < 3; ++a) {
            n[a] = thin[a] ? 1 : (int)std::max(1.0, std::min(1e6, ceil(ext[a] / h)));
            tot *= n[a];
        }
        if (tot <= cap)
            break;
        h *= 1.1;
    }
}

// ---------------------------------------------------------------------------
// Per-item arithmetic of the grid build (one cell, record or macro block)
// ---------------------------------------------------------------------------

// Record i of the grid, for triangle `id`: (v0, e1, e2) padded to 48 bytes with
// the id's bits in the padding.
LEM_DEV void make_rec(const float *__restrict__ tris, int32_t id, float4 out[3])
{
    Tri t = make_tri(tris + 9 * (int64_t)id);
    out[0] = make_float4(t.v0x, t.v0y, t.v0z, t.e1x);
    out[1] = make_float4(t.e1y, t.e1z, t.e2x, t.e2y);
    out[2] = make_float4(t.e2z, __builtin_bit_cast(float, id), 0.0f, 0.0f);
}

// Tight box of cell c's contents: the union of its triangles' eps-inflated
// boxes, clipped to the eps-inflated cell; an empty cell gets an inverted box.
// The cell's record range rides in the w lanes (start, end), so one 32-byte
// read gives the walk both the box and the range.
LEM_DEV void cell_box_of(const float *__restrict__ tris, const int32_t *__restrict__ cell_start,
                         const int32_t *__restrict__ ids, const GridDev &G, int64_t c,
                         float4 &out_lo, float4 &out_hi)
{
    const float INF = __builtin_huge_valf();
    float lo[3] = {INF, INF, INF}, hi[3] = {-INF, -INF, -INF};
    for (int32_t j = cell_start[c]; j < cell_start[c + 1]; ++j) {
        const float *t = tris + 9 * (int64_t)ids[j];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], fminf(t[a], fminf(t[3 + a], t[6 + a])) - G.eps);
            hi[a] = fmaxf(hi[a], fmaxf(t[a], fmaxf(t[3 + a], t[6 + a])) + G.eps);
        }
    }
    const int ix = (int)(c % G.nx), iy = (int)((c / G.nx) % G.ny), iz = (int)(c / ((int64_t)G.nx * G.ny));
    const float cl[3] = {G.lox + G.cx * ix - G.eps, G.loy + G.cy * iy - G.eps,
                         G.loz + G.cz * iz - G.eps};
    const float ch[3] = {G.lox + G.cx * (ix + 1) + G.eps, G.loy + G.cy * (iy + 1) + G.eps,
                         G.loz + G.cz * (iz + 1) + G.eps};
    if (cell_start[c + 1] > cell_start[c]) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = fmaxf(lo[a], cl[a]);
            hi[a] = fminf(hi[a], ch[a]);
        }
    }
    out_lo = make_float4(lo[0], lo[1], lo[2], __builtin_bit_cast(float, cell_start[c]));
    out_hi = make_float4(hi[0], hi[1], hi[2], __builtin_bit_cast(float, cell_start[c + 1]));
}

// Box of macro block m: the union of its cells' boxes (inverted if empty).
LEM_DEV void macro_box_of(const GridDev &G, const float4 *__restrict__ box, int64_t m,
                          float4 &out_lo, float4 &out_hi)
{
    const int jx = (int)(m % G.mx), jy = (int)((m / G.mx) % G.my), jz = (int)(m / ((int64_t)G.mx * G.my));
    const int B = 1 << G.mshift;
    const float INF = __builtin_huge_valf();
    float4 lo = make_float4(INF, INF, INF, 0.0f), hi = make_float4(-INF, -INF, -INF, 0.0f);
    for (int z = jz * B; z < min((jz + 1) * B, G.nz); ++z)
        for (int y = jy * B; y < min((jy + 1) * B, G.ny); ++y)
            for (int x = jx * B; x < min((jx + 1) * B, G.nx); ++x) {
                const int64_t c = ((int64_t)z * G.ny + y) * G.nx + x;
                const float4 a = box[2 * c], b = box[2 * c + 1];
                if (a.x <= b.x) {
                    lo.x = fminf(lo.x, a.x); lo.y = fminf(lo.y, a.y); lo.z = fminf(lo.z, a.z);
                    hi.x = fmaxf(hi.x, b.x); hi.y = fmaxf(hi.y, b.y); hi.z = fmaxf(hi.z, b.z);
                }
            }
    out_lo = lo;
    out_hi = hi;
}

// Macro box m quantised to 8 bytes: per axis, lo and hi as steps of
// (extent + 2 eps) / 255 from the block's origin minus eps, rounded outwards by
// a full step and clamped to the block widened by eps on each side (x: lo,
// y: lo, z: lo, x: hi | y: hi, z: hi); an empty block stores lo > hi.  Decoded
// with the same float32 origin and step (macro_origin of G.ql*, G.q*), the box
// covers the part of the union box within eps of the block -- every hit point
// the walk tests while it is in the block, also one on a block face that the
// walk's rounding assigns to this block.
LEM_DEV uint2 macro_quant_of(const GridDev &G, const float4 *__restrict__ box, int64_t m)
{
    const int jx = (int)(m % G.mx), jy = (int)((m / G.mx) % G.my), jz = (int)(m / ((int64_t)G.mx * G.my));
    const float4 lo = box[2 * (G.ncells + m)], hi = box[2 * (G.ncells + m) + 1];
    if (!(lo.x <= hi.x))
        return make_uint2(0x00ffffffu, 0u);  // lo = 255 > hi = 0 on every axis
    const float o[3] = {macro_origin(G.qlx, G.cx, jx, G.mshift),
                        macro_origin(G.qly, G.cy, jy, G.mshift),
                        macro_origin(G.qlz, G.cz, jz, G.mshift)};
    const float st[3] = {G.qx, G.qy, G.qz};
    const float l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z};
    unsigned int ql[3], qh[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double fl = floor(((double)l[a] - (double)o[a]) / (double)st[a]) - 1.0;
        const double fh = ceil(((double)h[a] - (double)o[a]) / (double)st[a]) + 1.0;
        ql[a] = (unsigned int)(fl < 0.0 ? 0.0 : (fl > 255.0 ? 255.0 : fl));
        qh[a] = (unsigned int)(fh < 0.0 ? 0.0 : (fh > 255.0 ? 255.0 : fh));
    }
    return make_uint2(ql[0] | (ql[1] << 8) | (ql[2] << 16) | (qh[0] << 24), qh[1] | (qh[2] << 8));
}

// Cell budget of the plan (the count array's capacity).
inline int64_t plan_cap(int64_t T, float cells_per_tri)
{
    double c = (double)T * (double)(cells_per_tri > 0 ? cells_per_tri : 2.0f) + 64.0;
    if (c > 1.0e9)
        c = 1.0e9;
    return (int64_t)c;
}

// Grid geometry from the scene bounds (the vertex-coordinate minima and maxima):
// eps, the origin and cell size, the resolution and the macro-block shift.
// Returns false for non-finite bounds.  ncells, eps, nmacro set; nrefs 0.
inline bool grid_geometry(const float blo[3], const float bhi[3], int64_t T, float cells_per_tri,
                          lem_grid_info *g)
{
    float lo[3], hi[3], ext[3];
    float span = 0.0f;
    for (int a = 0; a < 3; ++a) {
        lo[a] = blo[a];
        hi[a] = bhi[a];
        if (!(std::isfinite(lo[a]) && std::isfinite(hi[a])))
            return false;
        span = std::max(span, std::max(fabsf(lo[a]), fabsf(hi[a])));
        span = std::max(span, hi[a] - lo[a]);
    }
    // Inflation: far above the float32 error of cell stepping at this scale.
    const float eps = std::max(1e-4f, 2e-6f * span);
    for (int a = 0; a < 3; ++a) {
        lo[a] -= 2 * eps;
        hi[a] += 2 * eps;
        ext[a] = hi[a] - lo[a];
    }
    int n[3];
    choose_resolution(ext, (double)T * (cells_per_tri > 0 ? cells_per_tri : 2.0f),
                      plan_cap(T, cells_per_tri), n);
    for (int a = 0; a < 3; ++a) {
        g->lo[a] = lo[a];
        g->n[a] = n[a];
        g->cell[a] = ext[a] / (float)n[a];
    }
    g->ncells = (int64_t)n[0] * n[1] * n[2];
    g->nrefs = 0;
    g->eps = eps;
    for (g->mshift = LEM_GRID_MSHIFT;; ++g->mshift) {
        const int64_t B = (int64_t)1 << g->mshift;
        const int64_t nm = ((n[0] + B - 1) / B) * ((n[1] + B - 1) / B) * ((n[2] + B - 1) / B);
        g->nmacro = (int32_t)nm;
        if (nm <= MACRO_LDS_MAX || g->mshift >= 6)
            break;
    }
    return true;
}

}  // namespace lem
