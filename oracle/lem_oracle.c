/*
 * lem_oracle.c -- CPU oracle for LyceanEM's raycast + frequency-domain hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is compared
 * against (and the CPU baseline that bench.py times beside the GPU).  Nothing in
 * the product library links, loads or calls it.
 *
 * It restates, in plain C, the reference arithmetic cited next to each function
 * (paths relative to the LyceanEM-Python source tree).  Build: oracle/Makefile
 * (gcc -O2 -ffp-contract=off -fopenmp), so every FP32 expression is evaluated
 * as written, in source order, with no fused multiply-add.
 *
 * Pinning (see tests/test_oracle_*.py): closed-form hits/misses, exact rational
 * geometry on tiny scenes, both predicate semantics against each other away from
 * their documented differences, the projector identity, single-pair closed forms,
 * linearity in the weights and the unblocked planar-array factor.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

/* rayfunctions.py:19  EPSILON = 1e-7 (a Python float, i.e. double) */
static const double NUMBA_EPS = 1e-7;
/* math_utility.cuh:9  constexpr float EPSILON = 0.000001 */
static const float NATIVE_EPS = 0.000001f;

/* ------------------------------------------------------------------------ */
/* Ray charging                                                              */
/* ------------------------------------------------------------------------ */

/* Ray from origin o to target t: length L = |t-o| and unit direction (t-o)/L.
 * math_functions.py:148-156 (calc_dv_norm, numba njit on float32 arrays: the
 * difference, squares, sum and sqrt stay float32 -- numba resolves a float32
 * array ** int literal to the float32 loop -- then a float32 division), used by
 * charge_rays_environment1Dv2 (rayfunctions.py:1790-1842).  The native twin
 * points_to_rays (raycasting.cuh:80-95) evaluates the same expression. */
ORC_API void orc_charge(const float *o, const float *t, int64_t n, float *d_out,
                        float *len_out)
{
    for (int64_t i = 0; i < n; ++i) {
        float dx = t[3 * i + 0] - o[3 * i + 0];
        float dy = t[3 * i + 1] - o[3 * i + 1];
        float dz = t[3 * i + 2] - o[3 * i + 2];
        float s = dx * dx + dy * dy;
        s = s + dz * dz;
        float L = sqrtf(s);
        len_out[i] = L;
        d_out[3 * i + 0] = dx / L;
        d_out[3 * i + 1] = dy / L;
        d_out[3 * i + 2] = dz / L;
    }
}

static inline void charge_one(const float *o, const float *t, float *d, float *L)
{
    float dx = t[0] - o[0];
    float dy = t[1] - o[1];
    float dz = t[2] - o[2];
    float s = dx * dx + dy * dy;
    s = s + dz * dz;
    float l = sqrtf(s);
    *L = l;
    d[0] = dx / l;
    d[1] = dy / l;
    d[2] = dz / l;
}

/* ------------------------------------------------------------------------ */
/* Moller-Trumbore predicates                                                */
/* ------------------------------------------------------------------------ */

/* numba `hit` (rayfunctions.py:37-104), two-sided.
 * Edges, cross and dot products are float32 (rayfunctions.py:22-33 dot/cross,
 * evaluated left to right).  F = 1.0/A promotes to float64 (a Python float
 * literal), so U, V, U+V and t are float64.  The accept window compares the
 * float64 t against the ray's float32 `dist` promoted to float64.
 * tri = (v0x,v0y,v0z, v1x,v1y,v1z, v2x,v2y,v2z) like base_types.triangle_t.
 * Returns 1 on an accepted hit and writes the float64 distance. */
ORC_API int orc_hit_numba(const float *o, const float *d, float dist, const float *tri,
                          double *t_out)
{
    float e1x = tri[3] - tri[0], e1y = tri[4] - tri[1], e1z = tri[5] - tri[2];
    float e2x = tri[6] - tri[0], e2y = tri[7] - tri[1], e2z = tri[8] - tri[2];
    /* pvec = cross(d, e2) */
    float px = d[1] * e2z - d[2] * e2y;
    float py = d[2] * e2x - d[0] * e2z;
    float pz = d[0] * e2y - d[1] * e2x;
    /* A = dot(pvec, e1) */
    float A = px * e1x + py * e1y;
    A = A + pz * e1z;
    if (-NUMBA_EPS < (double)A && (double)A < NUMBA_EPS)
        return 0;
    float tx = o[0] - tri[0], ty = o[1] - tri[1], tz = o[2] - tri[2];
    double F = 1.0 / (double)A;
    float du = tx * px + ty * py;
    du = du + tz * pz;
    double U = F * (double)du;
    if (U < 0.0 || U > 1.0)
        return 0;
    /* q = cross(tvec, e1) */
    float qx = ty * e1z - tz * e1y;
    float qy = tz * e1x - tx * e1z;
    float qz = tx * e1y - ty * e1x;
    float dv = d[0] * qx + d[1] * qy;
    dv = dv + d[2] * qz;
    double V = F * (double)dv;
    if (V < 0.0 || (U + V) > 1.0)
        return 0;
    float dt = e2x * qx + e2y * qy;
    dt = dt + e2z * qz;
    double t = F * (double)dt;
    if (t > 2.0 * NUMBA_EPS && t < (double)dist - 2.0 * NUMBA_EPS) {
        *t_out = t;
        return 1;
    }
    return 0;
}

/* Native CUDA `hit` (raycasting.cuh:19-68): back-face culled (A<0 -> miss,
 * :31-34), EPSILON = 1e-6f, everything float32; F = 1.0/A is a double division
 * narrowed to float (:39).  helper_math dot/cross evaluate left to right. */
ORC_API int orc_hit_native(const float *o, const float *d, float len, const float *tri)
{
    float e1x = tri[3] - tri[0], e1y = tri[4] - tri[1], e1z = tri[5] - tri[2];
    float e2x = tri[6] - tri[0], e2y = tri[7] - tri[1], e2z = tri[8] - tri[2];
    float px = d[1] * e2z - d[2] * e2y;
    float py = d[2] * e2x - d[0] * e2z;
    float pz = d[0] * e2y - d[1] * e2x;
    /* dot(edge1, pvec) */
    float A = e1x * px + e1y * py;
    A = A + e1z * pz;
    if (fabsf(A) < NATIVE_EPS)
        return 0;
    if (A < 0)
        return 0;
    float tx = o[0] - tri[0], ty = o[1] - tri[1], tz = o[2] - tri[2];
    float F = (float)(1.0 / (double)A);
    float du = tx * px + ty * py;
    du = du + tz * pz;
    float U = F * du;
    if (U < 0.0 || U > 1.0)
        return 0;
    float qx = ty * e1z - tz * e1y;
    float qy = tz * e1x - tx * e1z;
    float qz = tx * e1y - ty * e1x;
    float dv = d[0] * qx + d[1] * qy;
    dv = dv + d[2] * qz;
    float V = F * dv;
    float UV = U + V;
    if (V < 0.0 || UV > 1.0)
        return 0;
    float dt = e2x * qx + e2y * qy;
    dt = dt + e2z * qz;
    float t = F * dt;
    float hi = len - 2 * NATIVE_EPS;
    if (t > 2 * NATIVE_EPS && t < hi)
        return 1;
    return 0;
}

/* kernel1Dv3 (rayfunctions.py:212-244) for one ray: every triangle in order;
 * an accepted hit with t < dist sets dist = (float32)t and intersect = True,
 * so later triangles see the shrunken window. */
static inline void raycast_one_numba(const float *o, const float *d, float dist0,
                                     const float *tris, int64_t T, uint8_t *hit,
                                     float *dist_out, int any_hit_only)
{
    float dist = dist0;
    uint8_t inter = 0;
    for (int64_t j = 0; j < T; ++j) {
        double t;
        if (orc_hit_numba(o, d, dist, tris + 9 * j, &t)) {
            if (t < (double)dist) {
                dist = (float)t;
                inter = 1;
                if (any_hit_only)
                    break;
            }
        }
    }
    *hit = inter;
    if (dist_out)
        *dist_out = dist;
}

/* launchRaycaster1Dv3 (rayfunctions.py:1951-2017) on an already charged ray
 * payload: rays is R x 7 float (ox,oy,oz,dx,dy,dz,dist) like base_types.ray_t. */
ORC_API void orc_raycast_rays(const float *rays, int64_t R, const float *tris, int64_t T,
                              uint8_t *intersect, float *dist_out, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < R; ++i) {
        const float *r = rays + 7 * i;
        raycast_one_numba(r, r + 3, r[6], tris, T, intersect + i, dist_out + i, 0);
    }
}

/* Dense S x K visibility.  semantics 0: numba charge + kernel1Dv3 over every
 * triangle (dist is the kernel's final `dist`); semantics 1: native culled
 * predicate, any hit (raycasting_brute_force.cuh:19-67 with a working copy-back).
 * skip_self (native flag == false): pairs with source index == sink index are
 * never tested (raycasting_accelerated.cuh:72). */
ORC_API void orc_visibility_dense(const float *src, int64_t S, const float *snk, int64_t K,
                                  const float *tris, int64_t T, int semantics,
                                  int any_hit_only, int skip_self, uint8_t *hit,
                                  float *dist, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t s = 0; s < S; ++s) {
        for (int64_t k = 0; k < K; ++k) {
            float d[3], L;
            charge_one(src + 3 * s, snk + 3 * k, d, &L);
            int64_t idx = s * K + k;
            if (skip_self && s == k) {
                hit[idx] = 0;
                if (dist)
                    dist[idx] = L;
                continue;
            }
            if (semantics == 0) {
                raycast_one_numba(src + 3 * s, d, L, tris, T, hit + idx,
                                  dist ? dist + idx : NULL, any_hit_only);
            } else {
                uint8_t h = 0;
                for (int64_t j = 0; j < T; ++j) {
                    if (orc_hit_native(src + 3 * s, d, L, tris + 9 * j)) {
                        h = 1;
                        break;
                    }
                }
                hit[idx] = h;
                if (dist)
                    dist[idx] = L;
            }
        }
    }
}

/* Sampled visibility: rays given by explicit (source row, sink row) pairs of
 * the dense problem, numba semantics.  Used for subsampled checks at sizes the
 * full oracle cannot finish. */
ORC_API void orc_visibility_sampled(const float *src, const float *snk, const int64_t *pairs,
                                    int64_t n, const float *tris, int64_t T,
                                    int any_hit_only, uint8_t *hit, float *dist,
                                    int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i) {
        const float *o = src + 3 * pairs[2 * i];
        const float *t = snk + 3 * pairs[2 * i + 1];
        float d[3], L;
        charge_one(o, t, d, &L);
        raycast_one_numba(o, d, L, tris, T, hit + i, dist ? dist + i : NULL, any_hit_only);
    }
}

/* ------------------------------------------------------------------------ */
/* Tile acceleration (the reference's cuda branch), as a CPU baseline         */
/* ------------------------------------------------------------------------ */

/* Tile_acceleration_structure (models/acceleration_structures.py:57-70) bins
 * each triangle into the 2-D tiles of its three vertices over the two
 * horizontal axes, plus the fourth tile of the 2 x 2 block when the vertices
 * fall in three different tiles (terrain_acceleration_build.cuh:57-112; bins
 * of one triangle must be adjacent, :114-133).  The reference takes x as the
 * vertical axis and tiles (y, z); `up` names the vertical axis here and the
 * tiles span the other two (a < b).  Tiles are `tile` wide from lo_a, lo_b.
 * The triangle ids of tile (i, j) are id[start[i * nb + j] .. start[... + 1]). */
typedef struct {
    int up, a, b;
    float lo_a, lo_b, tile;
    int na, nb;
    float bottom, top;  /* extent of the structure along `up` */
    int64_t *start;
    int32_t *id;
    const float *tris;
    int64_t T;
} orc_tiles;

static inline int tile_coord(float p, float lo, float tile) { return (int)floorf((p - lo) / tile); }

ORC_API void *orc_tiles_build(const float *tris, int64_t T, int up, float lo_a, float lo_b,
                              float tile, int na, int nb, float bottom, float top)
{
    orc_tiles *g = (orc_tiles *)calloc(1, sizeof(orc_tiles));
    g->up = up;
    g->a = up == 0 ? 1 : 0;
    g->b = up == 2 ? 1 : 2;
    g->lo_a = lo_a;
    g->lo_b = lo_b;
    g->tile = tile;
    g->na = na;
    g->nb = nb;
    g->bottom = bottom;
    g->top = top;
    g->tris = tris;
    g->T = T;
    const int64_t nt = (int64_t)na * nb;
    int64_t *cnt = (int64_t *)calloc((size_t)nt + 1, sizeof(int64_t));
    int32_t *bins = (int32_t *)malloc(sizeof(int32_t) * 4 * (size_t)(T > 0 ? T : 1));
    for (int64_t t = 0; t < T; ++t) {
        int ia[3], ib[3];
        for (int v = 0; v < 3; ++v) {
            ia[v] = tile_coord(tris[9 * t + 3 * v + g->a], lo_a, tile);
            ib[v] = tile_coord(tris[9 * t + 3 * v + g->b], lo_b, tile);
        }
        int n = 0;
        int32_t *bt = bins + 4 * t;
        for (int v = 0; v < 3; ++v) {
            int c = ia[v] * nb + ib[v];
            int dup = 0;
            for (int u = 0; u < n; ++u)
                dup |= bt[u] == c;
            if (!dup)
                bt[n++] = c;
        }
        if (n == 3) {
            /* the corner that completes the 2 x 2 block: on each axis the
             * coordinate that occurs once among the three tiles */
            int fa = ia[0] == ia[1] ? ia[2] : (ia[0] == ia[2] ? ia[1] : ia[0]);
            int fb = ib[0] == ib[1] ? ib[2] : (ib[0] == ib[2] ? ib[1] : ib[0]);
            bt[n++] = fa * nb + fb;
        }
        for (int u = n; u < 4; ++u)
            bt[u] = -1;
        for (int u = 0; u < n; ++u)
            if (bt[u] >= 0 && bt[u] < nt)
                cnt[bt[u] + 1]++;
    }
    for (int64_t c = 0; c < nt; ++c)
        cnt[c + 1] += cnt[c];
    g->start = cnt;
    g->id = (int32_t *)malloc(sizeof(int32_t) * (size_t)(cnt[nt] > 0 ? cnt[nt] : 1));
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nt > 0 ? nt : 1));
    memcpy(fill, cnt, sizeof(int64_t) * (size_t)nt);
    for (int64_t t = 0; t < T; ++t)
        for (int u = 0; u < 4; ++u) {
            int32_t c = bins[4 * t + u];
            if (c >= 0 && c < nt)
                g->id[fill[c]++] = (int32_t)t;
        }
    free(fill);
    free(bins);
    return g;
}

ORC_API int64_t orc_tiles_refs(const void *h)
{
    const orc_tiles *g = (const orc_tiles *)h;
    return g->start[(int64_t)g->na * g->nb];
}

ORC_API void orc_tiles_free(void *h)
{
    orc_tiles *g = (orc_tiles *)h;
    if (!g)
        return;
    free(g->start);
    free(g->id);
    free(g);
}

/* ray_plane_intersection (raycasting_accelerated.cuh:20-26) with the plane
 * up = level, float32: the point o - d / t * dir with d = o.up - level, t = dir.up */
static inline void plane_point(const orc_tiles *g, const float *o, const float *dir, float level,
                               float *pa, float *pb)
{
    const float d = o[g->up] - level;
    const float t = dir[g->up];
    *pa = o[g->a] - d / t * dir[g->a];
    *pb = o[g->b] - d / t * dir[g->b];
}

/* One ray of the tile DDA (raycasting_accelerated.cuh:106-268): the tiles
 * spanned by where the segment enters and leaves the structure's slab along
 * `up` -- the whole rectangle between them, row by row in the direction of
 * travel -- are scanned and every triangle binned there is tested until one
 * hits.  A segment entirely above or below the structure tests nothing.
 * sem 0: numba predicate with the ray's length as window; 1: native culled. */
static int tiles_cast(const orc_tiles *g, const float *o, const float *e, int sem)
{
    float dir[3], L;
    charge_one(o, e, dir, &L);
    const int up = g->up;
    const int o_above = o[up] > g->top, o_below = o[up] < g->bottom;
    const int o_in = !o_above && !o_below;
    const int e_above = e[up] > g->top, e_below = e[up] < g->bottom;
    const int e_in = !e_above && !e_below;
    float sa, sb, da, db;
    if (o[up] > e[up]) {
        if (o_above && e_below) {
            plane_point(g, o, dir, g->top, &sa, &sb);
            plane_point(g, o, dir, g->bottom, &da, &db);
        } else if (o_above && e_in) {
            plane_point(g, o, dir, g->top, &sa, &sb);
            da = e[g->a];
            db = e[g->b];
        } else if (e_below && o_in) {
            sa = o[g->a];
            sb = o[g->b];
            plane_point(g, o, dir, g->bottom, &da, &db);
        } else if (e_in && o_in) {
            sa = o[g->a];
            sb = o[g->b];
            da = e[g->a];
            db = e[g->b];
        } else {
            return 0;
        }
    } else {
        if (o_below && e_above) {
            plane_point(g, o, dir, g->bottom, &sa, &sb);
            plane_point(g, o, dir, g->top, &da, &db);
        } else if (o_below && e_in) {
            plane_point(g, o, dir, g->bottom, &sa, &sb);
            da = e[g->a];
            db = e[g->b];
        } else if (e_above && o_in) {
            sa = o[g->a];
            sb = o[g->b];
            plane_point(g, o, dir, g->top, &da, &db);
        } else if (e_in && o_in) {
            sa = o[g->a];
            sb = o[g->b];
            da = e[g->a];
            db = e[g->b];
        } else {
            return 0;
        }
    }
    const int ta = tile_coord(sa, g->lo_a, g->tile), tb = tile_coord(sb, g->lo_b, g->tile);
    const int ua = tile_coord(da, g->lo_a, g->tile), ub = tile_coord(db, g->lo_b, g->tile);
    /* step_direction (:44-48): the sign of the ray's direction on each axis */
    const int step_a = dir[g->a] < 0 ? -1 : 1, step_b = dir[g->b] < 0 ? -1 : 1;
    const int ra = abs(ua - ta) + 1, rb = abs(ub - tb) + 1;
    for (int i = 0; i < ra; ++i)
        for (int j = 0; j < rb; ++j) {
            const int ca = ta + i * step_a, cb = tb + j * step_b;
            if (ca < 0 || ca >= g->na || cb < 0 || cb >= g->nb)
                continue;
            const int64_t c = (int64_t)ca * g->nb + cb;
            for (int64_t q = g->start[c]; q < g->start[c + 1]; ++q) {
                const float *tri = g->tris + 9 * (int64_t)g->id[q];
                double t;
                if (sem == 1 ? orc_hit_native(o, dir, L, tri) : orc_hit_numba(o, dir, L, tri, &t))
                    return 1;
            }
        }
    return 0;
}

/* Any-hit masks of explicit (source row, sink row) pairs through the tiles. */
ORC_API void orc_tiles_cast(const void *h, const float *src, const float *snk,
                            const int64_t *pairs, int64_t n, int sem, uint8_t *hit, int nthreads)
{
    const orc_tiles *g = (const orc_tiles *)h;
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i)
        hit[i] = (uint8_t)tiles_cast(g, src + 3 * pairs[2 * i], snk + 3 * pairs[2 * i + 1], sem);
}

/* ------------------------------------------------------------------------ */
/* Fields (float64)                                                          */
/* ------------------------------------------------------------------------ */

typedef struct {
    double re, im;
} cplx;

static inline cplx cmul(cplx a, cplx b)
{
    cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}

/* sourcelaunchtransformGPU (empropagation.py:178-249), native twin ray_launch
 * (em.cuh:18-61): build u = normalize(axis x d) with the axis whose cross norm
 * is largest (ties as in the reference's if/elif chain), v = normalize(d x u),
 * and return (E.u)u + (E.v)v.  Float64 throughout. */
static void launch_transform(cplx E[3], const double dir[3])
{
    double axes[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double nrm[3];
    double cr[3][3];
    for (int a = 0; a < 3; ++a) {
        const double *x = axes[a];
        cr[a][0] = x[1] * dir[2] - x[2] * dir[1];
        cr[a][1] = x[2] * dir[0] - x[0] * dir[2];
        cr[a][2] = x[0] * dir[1] - x[1] * dir[0];
        nrm[a] = sqrt(cr[a][0] * cr[a][0] + cr[a][1] * cr[a][1] + cr[a][2] * cr[a][2]);
    }
    int pick = -1;
    if (fabs(nrm[0]) > fabs(nrm[1]) && fabs(nrm[0]) > fabs(nrm[2]))
        pick = 0;
    else if (fabs(nrm[1]) >= fabs(nrm[0]) && fabs(nrm[1]) > fabs(nrm[2]))
        pick = 1;
    else if (fabs(nrm[2]) >= fabs(nrm[0]) && fabs(nrm[2]) >= fabs(nrm[1]))
        pick = 2;
    double u[3] = {0, 0, 0}, v[3];
    if (pick >= 0) {
        for (int i = 0; i < 3; ++i)
            u[i] = cr[pick][i] / nrm[pick];
    }
    /* v = normalize(cross(dir, u)) */
    v[0] = dir[1] * u[2] - dir[2] * u[1];
    v[1] = dir[2] * u[0] - dir[0] * u[2];
    v[2] = dir[0] * u[1] - dir[1] * u[0];
    double vn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int i = 0; i < 3; ++i)
        v[i] /= vn;
    cplx eu = {0, 0}, ev = {0, 0};
    for (int i = 0; i < 3; ++i) {
        eu.re += E[i].re * u[i];
        eu.im += E[i].im * u[i];
        ev.re += E[i].re * v[i];
        ev.im += E[i].im * v[i];
    }
    for (int i = 0; i < 3; ++i) {
        E[i].re = eu.re * u[i] + ev.re * v[i];
        E[i].im = eu.im * u[i] + ev.im * v[i];
    }
}

/* lossy_propagation (empropagation.py:768-801), native twin em_wave
 * (em.cuh:64-120): r = |p2-p1| (optionally rounded to float32 as the numba
 * kernel's float32 local does, :771-772), d = (p2-p1)/r,
 * G = exp(-(alpha+j beta) r)/r, dG = (-(alpha+j beta) - 1/r) G,
 * loss = -(1/2pi) dG (d . n2) with the end point's raw normal. */
static cplx lossy(const float *p1, const float *p2, const float *n2, double alpha,
                  double beta, int r_f32, double dir[3])
{
    double dx = (double)p2[0] - (double)p1[0];
    double dy = (double)p2[1] - (double)p1[1];
    double dz = (double)p2[2] - (double)p1[2];
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    if (r_f32)
        r = (double)(float)r;
    dir[0] = dx / r;
    dir[1] = dy / r;
    dir[2] = dz / r;
    double cosang = dir[0] * (double)n2[0] + dir[1] * (double)n2[1] + dir[2] * (double)n2[2];
    double front = -(1.0 / (2.0 * M_PI));
    double inv = 1.0 / r;
    double amp = exp(-alpha * r) * inv;
    cplx G = {amp * cos(-beta * r), amp * sin(-beta * r)};
    cplx a = {-alpha - inv, -beta};
    cplx dG = cmul(a, G);
    cplx loss = {front * dG.re * cosang, front * dG.im * cosang};
    return loss;
}

static inline cplx ldc(const float *w, int64_t i, int c)
{
    cplx z = {(double)w[6 * i + 2 * c], (double)w[6 * i + 2 * c + 1]};
    return z;
}

/* One line-of-sight path source s -> sink k of freqdomainkernal
 * (empropagation.py:1083-1265): E = w_s launched along d_sk, times the sink
 * weight (:1168-1179), optionally projected onto the sink's tangent plane as the
 * kernel does for LOS rows of a 3-column index (:1150-1161 with i == 1), times
 * the leg loss. */
static void los_pair(const float *ps, const float *ws, int64_t s, const float *pk,
                     const float *nk, const float *wk, int64_t k, double alpha, double beta,
                     int flags, cplx out[3])
{
    double dir[3];
    cplx loss = lossy(ps + 3 * s, pk + 3 * k, nk + 3 * k, alpha, beta, flags & 4, dir);
    cplx E[3];
    for (int c = 0; c < 3; ++c)
        E[c] = ldc(ws, s, c);
    launch_transform(E, dir);
    if (wk) {
        for (int c = 0; c < 3; ++c)
            E[c] = cmul(E[c], ldc(wk, k, c));
    }
    if (flags & 1) {
        double n[3] = {nk[3 * k], nk[3 * k + 1], nk[3 * k + 2]};
        launch_transform(E, n);
    }
    for (int c = 0; c < 3; ++c)
        out[c] = cmul(E[c], loss);
}

/* Dense line-of-sight network in float64.
 * src_pos S x 3 float, src_w S x 3 complex64 (interleaved re,im), sink_pos /
 * sink_nrm K x 3 float, sink_w K x 3 complex64 or NULL (weight 1 as set by
 * frequency_domain.py:398-400), hit S x K (1 = blocked) or NULL.
 * flags: 1 = project onto the sink tangent plane (3-column index quirk),
 *        2 = elements (write the S x K x 3 network instead of the source sum),
 *        4 = round path lengths to float32 like the numba kernel.
 * out: K x 3 complex128 (sum over sources, Kahan-compensated long double) or
 * S x K x 3 complex128 when flags & 2. */
ORC_API void orc_field_los(const float *src_pos, const float *src_w, int64_t S,
                           const float *snk_pos, const float *snk_nrm, const float *snk_w,
                           int64_t K, const uint8_t *hit, double alpha, double beta, int flags,
                           double *out, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#endif
    if (flags & 2) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t s = 0; s < S; ++s) {
            for (int64_t k = 0; k < K; ++k) {
                double *o = out + 6 * (s * K + k);
                if (hit && hit[s * K + k]) {
                    for (int i = 0; i < 6; ++i)
                        o[i] = 0.0;
                    continue;
                }
                cplx c3[3];
                los_pair(src_pos, src_w, s, snk_pos, snk_nrm, snk_w, k, alpha, beta, flags, c3);
                for (int c = 0; c < 3; ++c) {
                    o[2 * c] = c3[c].re;
                    o[2 * c + 1] = c3[c].im;
                }
            }
        }
        return;
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t k = 0; k < K; ++k) {
        long double acc[6] = {0, 0, 0, 0, 0, 0};
        long double cmp[6] = {0, 0, 0, 0, 0, 0};
        for (int64_t s = 0; s < S; ++s) {
            if (hit && hit[s * K + k])
                continue;
            cplx c3[3];
            los_pair(src_pos, src_w, s, snk_pos, snk_nrm, snk_w, k, alpha, beta, flags, c3);
            double v[6] = {c3[0].re, c3[0].im, c3[1].re, c3[1].im, c3[2].re, c3[2].im};
            for (int i = 0; i < 6; ++i) {
                long double y = (long double)v[i] - cmp[i];
                long double t = acc[i] + y;
                cmp[i] = (t - acc[i]) - y;
                acc[i] = t;
            }
        }
        for (int i = 0; i < 6; ++i)
            out[6 * k + i] = (double)acc[i];
    }
}

/* freqdomainkernal (empropagation.py:1083-1265) over an explicit path list, the
 * network EMGPUFreqDomain returns (:2197-2410).  Float64, paths summed in index
 * order (the reference adds with float64 atomics in an unspecified order).
 * pos / nrm: N x 3 float; w: N x 3 complex64 (interleaved) -- the scattering_t
 * table (base_types.py:89-119).  eta (N doubles, or NULL when every point is
 * electric): eta[i] > 0 marks a magnetic point (Electric = False) of impedance
 * sqrt(Re permeability / Re permittivity); a path starting there launches
 * (w x d) / eta, d the unit direction to its next point
 * (empropagation.py:1111-1146), instead of w.
 * index: V x ncol int64, 1-based, 0 = no point (a padded LOS row).
 * The (source, sink) slot of each row comes from targettingindex
 * (empropagation.py:3394-3429): LOS rows -> (col0, col1), others -> (col0, last).
 * out: S x K x 3 complex128, zeroed here. */
ORC_API void orc_field_paths(const float *pos, const float *nrm, const float *w,
                             const double *eta, int64_t N_points, const int64_t *index, int64_t V,
                             int ncol, int64_t S, int64_t K, double alpha, double beta, int r_f32,
                             double *out)
{
    memset(out, 0, sizeof(double) * 6 * (size_t)(S * K));
    for (int64_t p = 0; p < V; ++p) {
        const int64_t *row = index + (int64_t)ncol * p;
        cplx E[3];
        int64_t last = row[1];
        for (int i = 0; i < ncol - 1; ++i) {
            /* a zero entry (no point) reads point -1, the last row of the table:
             * numba wraps the negative index network_index - 1
             * (empropagation.py:1147-1156) */
            int64_t a = row[i] > 0 ? row[i] - 1 : N_points - 1;
            int64_t b = row[i + 1];
            if (i == 0) {
                for (int c = 0; c < 3; ++c)
                    E[c] = ldc(w, a, c);
                if (eta && eta[a] > 0 && b != 0) {
                    const float *pa = pos + 3 * a, *pb = pos + 3 * (b - 1);
                    double d[3] = {(double)pb[0] - pa[0], (double)pb[1] - pa[1],
                                   (double)pb[2] - pa[2]};
                    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                    for (int c = 0; c < 3; ++c)
                        d[c] /= r;
                    cplx m[3];
                    for (int c = 0; c < 3; ++c) {
                        /* (w x d)_c = w_{c+1} d_{c+2} - w_{c+2} d_{c+1} */
                        const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
                        m[c].re = (E[c1].re * d[c2] - E[c2].re * d[c1]) / eta[a];
                        m[c].im = (E[c1].im * d[c2] - E[c2].im * d[c1]) / eta[a];
                    }
                    for (int c = 0; c < 3; ++c)
                        E[c] = m[c];
                }
            } else {
                double n[3] = {nrm[3 * a], nrm[3 * a + 1], nrm[3 * a + 2]};
                launch_transform(E, n);
            }
            if (b != 0) {
                double dir[3];
                const float *pa = pos + 3 * a, *pb = pos + 3 * (b - 1);
                double dx = (double)pb[0] - pa[0], dy = (double)pb[1] - pa[1],
                       dz = (double)pb[2] - pa[2];
                double r = sqrt(dx * dx + dy * dy + dz * dz);
                dir[0] = dx / r;
                dir[1] = dy / r;
                dir[2] = dz / r;
                launch_transform(E, dir);
                for (int c = 0; c < 3; ++c)
                    E[c] = cmul(E[c], ldc(w, b - 1, c));
                last = b;
            }
        }
        double dir[3];
        cplx loss = lossy(pos + 3 * (row[0] - 1), pos + 3 * (row[1] - 1), nrm + 3 * (row[1] - 1),
                          alpha, beta, r_f32, dir);
        for (int i = 1; i < ncol - 1; ++i) {
            if (row[i + 1] != 0) {
                int64_t a = row[i] > 0 ? row[i] - 1 : N_points - 1;
                cplx l2 = lossy(pos + 3 * a, pos + 3 * (row[i + 1] - 1),
                                nrm + 3 * (row[i + 1] - 1), alpha, beta, r_f32, dir);
                loss = cmul(loss, l2);
            }
        }
        int64_t s = row[0] - 1;
        int64_t k = last - 1 - S;
        if (s < 0 || s >= S || k < 0 || k >= K)
            continue;
        double *o = out + 6 * (s * K + k);
        for (int c = 0; c < 3; ++c) {
            cplx v = cmul(E[c], loss);
            o[2 * c] += v.re;
            o[2 * c + 1] += v.im;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Time domain (timedomainkernal, empropagation.py:1345-1612)                */
/* ------------------------------------------------------------------------ */

/* Python float floor division a // b as CPython's float_floor_div computes it
 * (numba's float floordiv follows it): fmod, then exact adjustment.  Used for
 * the start sample time_offset // time_sep (empropagation.py:1503-1506). */
ORC_API double orc_py_floordiv(double a, double b)
{
    double mod = fmod(a, b);
    double div = (a - mod) / b;
    if (mod != 0.0 && ((b < 0.0) != (mod < 0.0)))
        div -= 1.0;
    if (div != 0.0) {
        double fl = floor(div);
        if (div - fl > 0.5)
            fl += 1.0;
        return fl;
    }
    return copysign(0.0, a / b);
}

/* point table index of a 1-based entry; 0 reads point -1 (numba wraps
 * negative indices to the end of the array) */
static int64_t orc_pt(int32_t e, int64_t N) { return e >= 1 ? (int64_t)e - 1 : N - 1; }

/* calc_sep (empropagation.py:82-93): float32 differences, float64 squares,
 * sum and square root */
static double orc_sep(const float *p, int64_t a, int64_t b)
{
    double dx = (double)(p[3 * b] - p[3 * a]), dy = (double)(p[3 * b + 1] - p[3 * a + 1]),
           dz = (double)(p[3 * b + 2] - p[3 * a + 2]);
    return sqrt(dx * dx + dy * dy + dz * dz);
}

/* lossy_propagation (empropagation.py:768-801) as a complex value:
 *   length  float32(calc_sep)            (the float32 local array, :771-772)
 *   d       float32(float32 diff / float64 norm)   (calc_dv into float32)
 *   dot     float32 dot of d and the end point's normal
 *   G       exp(-(alpha + j beta) len) / len       (complex128)
 *   dG      (-(alpha + j beta) - complex64(1/len)) G
 *   loss    -(1/2pi) dG dot */
static void orc_lossy(const float *p, const float *n, int64_t a, int64_t b, double alpha,
                      double beta, double *re, double *im)
{
    float dx = p[3 * b] - p[3 * a], dy = p[3 * b + 1] - p[3 * a + 1], dz = p[3 * b + 2] - p[3 * a + 2];
    double nrm = sqrt((double)dx * dx + (double)dy * dy + (double)dz * dz);
    float len = (float)nrm;
    float ux = (float)((double)dx / nrm), uy = (float)((double)dy / nrm), uz = (float)((double)dz / nrm);
    float dot = ux * n[3 * b] + uy * n[3 * b + 1];
    dot = dot + uz * n[3 * b + 2];
    double inv = 1.0 / (double)len;
    double ea = exp(-alpha * (double)len);
    double gr = ea * cos(-beta * (double)len) * inv, gi = ea * sin(-beta * (double)len) * inv;
    double ar = -alpha - (double)(float)inv, ai = -beta;
    double dgr = ar * gr - ai * gi, dgi = ar * gi + ai * gr;
    double f = -(1.0 / (2.0 * M_PI)) * (double)dot;
    *re = f * dgr;
    *im = f * dgi;
}

/* sourcelaunchtransformGPU (empropagation.py:178-249) applied to a real
 * vector: the components along u and v, the normalised axis x d / |x d| (the
 * global axis with the largest |axis x d|) and d x u / |d x u|, for the
 * normalised direction d. */
static void orc_launch(const double d_in[3], double v[3])
{
    double m = sqrt(d_in[0] * d_in[0] + d_in[1] * d_in[1] + d_in[2] * d_in[2]);
    double d[3] = {d_in[0] / m, d_in[1] / m, d_in[2] / m};
    double ax[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double cn[3];
    for (int k = 0; k < 3; ++k) {
        double c0 = ax[k][1] * d[2] - ax[k][2] * d[1], c1 = ax[k][2] * d[0] - ax[k][0] * d[2],
               c2 = ax[k][0] * d[1] - ax[k][1] * d[0];
        cn[k] = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    }
    int k = 2;
    if (fabs(cn[0]) > fabs(cn[1]) && fabs(cn[0]) > fabs(cn[2]))
        k = 0;
    else if (fabs(cn[1]) >= fabs(cn[0]) && fabs(cn[1]) > fabs(cn[2]))
        k = 1;
    double u[3] = {ax[k][1] * d[2] - ax[k][2] * d[1], ax[k][2] * d[0] - ax[k][0] * d[2],
                   ax[k][0] * d[1] - ax[k][1] * d[0]};
    double un = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    u[0] /= un; u[1] /= un; u[2] /= un;
    double w[3] = {d[1] * u[2] - d[2] * u[1], d[2] * u[0] - d[0] * u[2], d[0] * u[1] - d[1] * u[0]};
    double wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    w[0] /= wn; w[1] /= wn; w[2] /= wn;
    double eu = v[0] * u[0] + v[1] * u[1] + v[2] * u[2];
    double ev = v[0] * w[0] + v[1] * w[1] + v[2] * w[2];
    for (int c = 0; c < 3; ++c)
        v[c] = eu * u[c] + ev * w[c];
}

/* Per path row (timedomainkernal, empropagation.py:1361-1491): ray_component
 * = Re(weight of the first point) -- for a magnetic first point (eta[p] > 0,
 * eta as for orc_field_paths) Re(w) x d / eta, d the unit direction of the
 * first leg (:1395-1419) -- then per leg i -> i+1: (i > 0) the launch
 * transform with point i's normal and the leg length added; if the next entry
 * is non-zero, the launch transform along the leg, times Re(next weight), the
 * delay increased by Im(next ex), and (i > 0) the loss product extended.
 * amp = ray_component * |loss product| (cmath.polar), arrival = length / c +
 * delay.  w: N x 3 complex64 as interleaved float pairs. */
ORC_API void orc_time_paths(const float *pos, const float *nrm, const float *w,
                            const double *eta, int64_t N, const int32_t *index, int64_t V,
                            int ncol, double alpha, double beta, double *amp, double *arrival)
{
    const double C = 299792458.0;
    for (int64_t v = 0; v < V; ++v) {
        const int32_t *row = index + v * ncol;
        double rc[3] = {0, 0, 0}, len = 0.0, delay = 0.0;
        for (int i = 0; i < ncol - 1; ++i) {
            int64_t a = orc_pt(row[i], N), b = orc_pt(row[i + 1], N);
            if (i == 0) {
                len = orc_sep(pos, a, b);
                delay += (double)w[6 * a + 1];
                rc[0] = w[6 * a];
                rc[1] = w[6 * a + 2];
                rc[2] = w[6 * a + 4];
                if (eta && eta[a] > 0) {
                    double d[3] = {(double)(pos[3 * b] - pos[3 * a]),
                                   (double)(pos[3 * b + 1] - pos[3 * a + 1]),
                                   (double)(pos[3 * b + 2] - pos[3 * a + 2])};
                    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                    for (int c = 0; c < 3; ++c)
                        d[c] /= r;
                    double m[3];
                    for (int c = 0; c < 3; ++c) {
                        const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
                        m[c] = (rc[c1] * d[c2] - rc[c2] * d[c1]) / eta[a];
                    }
                    for (int c = 0; c < 3; ++c)
                        rc[c] = m[c];
                }
            } else {
                double nn[3] = {nrm[3 * a], nrm[3 * a + 1], nrm[3 * a + 2]};
                orc_launch(nn, rc);
                len = orc_sep(pos, a, b) + len;
            }
            if (row[i + 1] != 0) {
                double dd[3] = {(double)(pos[3 * b] - pos[3 * a]), (double)(pos[3 * b + 1] - pos[3 * a + 1]),
                                (double)(pos[3 * b + 2] - pos[3 * a + 2])};
                orc_launch(dd, rc);
                rc[0] *= (double)w[6 * b];
                rc[1] *= (double)w[6 * b + 2];
                rc[2] *= (double)w[6 * b + 4];
                delay += (double)w[6 * b + 1];
            }
        }
        double lr, li;
        orc_lossy(pos, nrm, orc_pt(row[0], N), orc_pt(row[1], N), alpha, beta, &lr, &li);
        for (int i = 1; i < ncol - 1; ++i) {
            if (row[i + 1] == 0)
                continue;
            double br, bi;
            orc_lossy(pos, nrm, orc_pt(row[i], N), orc_pt(row[i + 1], N), alpha, beta, &br, &bi);
            double nr = lr * br - li * bi, ni = lr * bi + li * br;
            lr = nr;
            li = ni;
        }
        double mag = hypot(lr, li);
        amp[3 * v] = rc[0] * mag;
        amp[3 * v + 1] = rc[1] * mag;
        amp[3 * v + 2] = rc[2] * mag;
        arrival[v] = len / C + delay;
    }
}

/* Placement (empropagation.py:1490-1600): wake = min(1.0, arrivals) (the
 * atomic minimum from its initial 1 s), start sample = (arrival - wake) //
 * (1 / fs), then excitation[j] * amp added at start + j for j < min(E,
 * N - start), into out[slot[v]][n][c] (nslots x N x 3, zero on entry), paths
 * in index order.  Writes *wake and bins[v]. */
ORC_API void orc_time_place(const double *amp, const double *arrival, int64_t V,
                            const int64_t *slot, const double *exc, int64_t E, double fs,
                            int64_t N, double *out, double *wake, int64_t *bins)
{
    double wk = 1.0;
    for (int64_t v = 0; v < V; ++v)
        if (arrival[v] < wk)
            wk = arrival[v];
    *wake = wk;
    const double sep = 1.0 / fs;
    for (int64_t v = 0; v < V; ++v) {
        double b = orc_py_floordiv(arrival[v] - wk, sep);
        int64_t s0 = (int64_t)b;
        bins[v] = s0;
        int64_t cnt = s0 + E <= N ? E : N - s0;
        for (int64_t j = 0; j < cnt; ++j)
            for (int c = 0; c < 3; ++c)
                out[(slot[v] * N + s0 + j) * 3 + c] += exc[j] * amp[3 * v + c];
    }
}
