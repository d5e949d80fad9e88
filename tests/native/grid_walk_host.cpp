// grid_walk_host.cpp -- CPU build of the uniform grid and its two-level DDA walk
// (lyceanem-python_amd/csrc/lem_grid_walk.h, the code the gfx950 traversal
// kernels run) checked ray by ray against the oracle's brute force over every
// triangle (oracle/lem_oracle.c: kernel1Dv3 with the numba predicate, and the
// native culled predicate).  Test infrastructure, built and run by
// tests/test_grid_walk_host.py, also under AddressSanitizer and
// UndefinedBehaviorSanitizer.
//
// The grid is built here by host loops from the same per-item functions the
// build kernels call (tri_cell_range, make_rec, cell_box_of, macro_box_of,
// macro_quant_of, grid_geometry); each cell lists its triangles by ascending id
// as k_sort_cells leaves them.  Every ray is cast the way each traversal kernel
// casts it:
//   global  -- dda_cast with the float macro boxes (k_visibility_grid);
//   lds     -- the quantised macro table (k_visibility_grid_lds for rays with no
//              zero direction component, k_visibility_grid_par for the others);
//   rev     -- any-hit walked from the sink end with the source -> sink ray
//              deciding (LEM_VIS_FROM_SINK), in both table forms.
// Prints one line per scene / grid / variant and exits 1 on any mask or
// distance mismatch.
#include "lem_host_shim.h"

#include "../../lyceanem-python_amd/csrc/lem_grid_walk.h"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

extern "C" void orc_visibility_dense(const float *src, int64_t S, const float *snk, int64_t K,
                                     const float *tris, int64_t T, int semantics,
                                     int any_hit_only, int skip_self, uint8_t *hit, float *dist,
                                     int nthreads);

using namespace lem;

namespace {

// ---- seeded generator (splitmix64) -----------------------------------------
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uni(double a, double b) { return a + (b - a) * (double)(next() >> 11) * 0x1.0p-53; }
    double gauss()
    {
        double u = uni(1e-12, 1.0), v = uni(0.0, 1.0);
        return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v);
    }
};

// ---- host grid ------------------------------------------------------------
struct HostGrid {
    lem_grid_info info;
    GridDev G;
    std::vector<int32_t> cell_start, ids;
    std::vector<float4> recs, box;   // box: 2 float4 per cell, then per macro block
    std::vector<uint32_t> qa;        // quantised macro table as the LDS holds it
    std::vector<uint16_t> qb;
    float4 axis[3];                  // the per-axis constants the LDS holds beside the table
};

// force_mshift < 0: the plan's own choice.
HostGrid build_grid(const std::vector<float> &tris, float cells_per_tri, int force_mshift)
{
    const int64_t T = (int64_t)tris.size() / 9;
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = __builtin_huge_valf();
        hi[a] = -__builtin_huge_valf();
    }
    for (int64_t v = 0; v < 3 * T; ++v)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], tris[3 * v + a]);
            hi[a] = std::max(hi[a], tris[3 * v + a]);
        }
    HostGrid H;
    std::memset(&H.info, 0, sizeof(H.info));
    if (!grid_geometry(lo, hi, T, cells_per_tri, &H.info)) {
        std::fprintf(stderr, "non-finite bounds\n");
        std::exit(2);
    }
    if (force_mshift >= 0) {
        H.info.mshift = force_mshift;
        const int64_t B = (int64_t)1 << force_mshift;
        const int *n = H.info.n;
        H.info.nmacro = (int32_t)(((n[0] + B - 1) / B) * ((n[1] + B - 1) / B) *
                                  ((n[2] + B - 1) / B));
    }
    H.G = to_dev(H.info);
    const GridDev &G = H.G;
    const int64_t nc = H.info.ncells;
    std::vector<int64_t> counts(nc, 0);
    for (int64_t t = 0; t < T; ++t) {
        int l[3], h[3];
        tri_cell_range(G, &tris[9 * t], l, h);
        for (int z = l[2]; z <= h[2]; ++z)
            for (int y = l[1]; y <= h[1]; ++y)
                for (int x = l[0]; x <= h[0]; ++x)
                    counts.at(((int64_t)z * G.ny + y) * G.nx + x) += 1;
    }
    H.cell_start.assign(nc + 1, 0);
    for (int64_t c = 0; c < nc; ++c)
        H.cell_start[c + 1] = H.cell_start[c] + (int32_t)counts[c];
    H.info.nrefs = H.cell_start[nc];
    H.ids.assign(H.info.nrefs, -1);
    std::vector<int32_t> cursor(H.cell_start.begin(), H.cell_start.end() - 1);
    for (int64_t t = 0; t < T; ++t) {  // triangle order: each cell's ids ascend
        int l[3], h[3];
        tri_cell_range(G, &tris[9 * t], l, h);
        for (int z = l[2]; z <= h[2]; ++z)
            for (int y = l[1]; y <= h[1]; ++y)
                for (int x = l[0]; x <= h[0]; ++x)
                    H.ids.at(cursor.at(((int64_t)z * G.ny + y) * G.nx + x)++) = (int32_t)t;
    }
    H.recs.resize(3 * H.info.nrefs);
    for (int64_t i = 0; i < H.info.nrefs; ++i)
        make_rec(tris.data(), H.ids[i], &H.recs[3 * i]);
    H.box.resize(2 * (nc + H.info.nmacro));
    for (int64_t c = 0; c < nc; ++c)
        cell_box_of(tris.data(), H.cell_start.data(), H.ids.data(), G, c, H.box[2 * c],
                    H.box[2 * c + 1]);
    for (int64_t m = 0; m < H.info.nmacro; ++m)
        macro_box_of(G, H.box.data(), m, H.box[2 * (nc + m)], H.box[2 * (nc + m) + 1]);
    for (int a = 0; a < 3; ++a)
        H.axis[a] = axis_entry(G, a);
    H.qa.resize(H.info.nmacro);
    H.qb.resize(H.info.nmacro);
    for (int64_t m = 0; m < H.info.nmacro; ++m) {
        const uint2 q = macro_quant_of(G, H.box.data(), m);
        H.qa[m] = q.x;
        H.qb[m] = (uint16_t)q.y;
    }
    return H;
}

// ---- the kernels' per-ray casts ---------------------------------------------
enum Variant { GLOBAL, LDS, REV_GLOBAL, REV_LDS };

bool cast(const HostGrid &H, Variant v, bool any, int sem, const float *s, const float *k,
          float &dist)
{
    const Ray r = charge(s[0], s[1], s[2], k[0], k[1], k[2]);
    dist = r.len;
    const GridDev &G = H.G;
    const int32_t *cs = H.cell_start.data();
    const float4 *rc = H.recs.data(), *bx = H.box.data();
    const MacroTab smq{H.qa.data(), H.qb.data(), H.axis};
    if (v == GLOBAL) {
        if (sem == 1)
            return dda_cast<true, 1>(r, dist, G, cs, rc, bx);
        return any ? dda_cast<true, 0>(r, dist, G, cs, rc, bx)
                   : dda_cast<false, 0>(r, dist, G, cs, rc, bx);
    }
    if (v == LDS) {
        const bool par = r.dx == 0.0f || r.dy == 0.0f || r.dz == 0.0f;
        if (sem == 1)
            return par ? dda_cast<true, 1, true, true>(r, dist, G, cs, rc, bx, nullptr, nullptr,
                                                        nullptr, smq)
                       : dda_cast<true, 1, true, false>(r, dist, G, cs, rc, bx, nullptr, nullptr,
                                                         nullptr, smq);
        if (any)
            return par ? dda_cast<true, 0, true, true>(r, dist, G, cs, rc, bx, nullptr, nullptr,
                                                        nullptr, smq)
                       : dda_cast<true, 0, true, false>(r, dist, G, cs, rc, bx, nullptr, nullptr,
                                                         nullptr, smq);
        return par ? dda_cast<false, 0, true, true>(r, dist, G, cs, rc, bx, nullptr, nullptr,
                                                     nullptr, smq)
                   : dda_cast<false, 0, true, false>(r, dist, G, cs, rc, bx, nullptr, nullptr,
                                                      nullptr, smq);
    }
    // walked from the sink end; the source -> sink ray decides (any-hit only)
    const Ray rw = charge(k[0], k[1], k[2], s[0], s[1], s[2]);
    float dw = rw.len;
    if (v == REV_GLOBAL)
        return sem == 1 ? dda_cast<true, 1>(rw, dw, G, cs, rc, bx, nullptr, nullptr, nullptr,
                                            MacroTab{nullptr, nullptr, nullptr}, &r)
                        : dda_cast<true, 0>(rw, dw, G, cs, rc, bx, nullptr, nullptr, nullptr,
                                            MacroTab{nullptr, nullptr, nullptr}, &r);
    const bool par = rw.dx == 0.0f || rw.dy == 0.0f || rw.dz == 0.0f;
    if (sem == 1)
        return par ? dda_cast<true, 1, true, true>(rw, dw, G, cs, rc, bx, nullptr, nullptr,
                                                    nullptr, smq, &r)
                   : dda_cast<true, 1, true, false>(rw, dw, G, cs, rc, bx, nullptr, nullptr,
                                                     nullptr, smq, &r);
    return par ? dda_cast<true, 0, true, true>(rw, dw, G, cs, rc, bx, nullptr, nullptr, nullptr,
                                                smq, &r)
               : dda_cast<true, 0, true, false>(rw, dw, G, cs, rc, bx, nullptr, nullptr, nullptr,
                                                 smq, &r);
}

struct Scene {
    std::string name;
    std::vector<float> tris, src, snk;
};

int64_t g_bad = 0;
int64_t g_rays = 0;

void check_scene(const Scene &sc, float cells_per_tri, int force_mshift)
{
    const int64_t T = (int64_t)sc.tris.size() / 9, S = (int64_t)sc.src.size() / 3,
                  K = (int64_t)sc.snk.size() / 3;
    const HostGrid H = build_grid(sc.tris, cells_per_tri, force_mshift);
    std::vector<uint8_t> hit(S * K), hit_n(S * K);
    std::vector<float> dist(S * K);
    orc_visibility_dense(sc.src.data(), S, sc.snk.data(), K, sc.tris.data(), T, 0, 0, 0,
                         hit.data(), dist.data(), 0);
    orc_visibility_dense(sc.src.data(), S, sc.snk.data(), K, sc.tris.data(), T, 1, 1, 0,
                         hit_n.data(), nullptr, 0);
    int64_t blocked = 0;
    for (auto h : hit)
        blocked += h;
    struct Case {
        const char *name;
        Variant v;
        bool any;
        int sem;
    } cases[] = {{"closest/global", GLOBAL, false, 0}, {"closest/lds", LDS, false, 0},
                 {"any/global", GLOBAL, true, 0},      {"any/lds", LDS, true, 0},
                 {"any/rev/global", REV_GLOBAL, true, 0}, {"any/rev/lds", REV_LDS, true, 0},
                 {"native/global", GLOBAL, true, 1},   {"native/lds", LDS, true, 1},
                 {"native/rev/lds", REV_LDS, true, 1}};
    for (const Case &c : cases) {
        int64_t bad_m = 0, bad_d = 0;
        for (int64_t s = 0; s < S; ++s)
            for (int64_t k = 0; k < K; ++k) {
                float d;
                const bool h = cast(H, c.v, c.any, c.sem, &sc.src[3 * s], &sc.snk[3 * k], d);
                const uint8_t ref = c.sem == 1 ? hit_n[s * K + k] : hit[s * K + k];
                bad_m += (uint8_t)h != ref;
                if (!c.any && std::memcmp(&d, &dist[s * K + k], 4) != 0)
                    ++bad_d;
            }
        g_bad += bad_m + bad_d;
        g_rays += S * K;
        std::printf("%-14s cells/tri %.0f mshift %d (%d x %d x %d cells, %lld refs): %-15s "
                    "%lld rays, blocked %.3f, %lld mask and %lld distance mismatches\n",
                    sc.name.c_str(), cells_per_tri, H.info.mshift, H.info.n[0], H.info.n[1],
                    H.info.n[2], (long long)H.info.nrefs, c.name, (long long)(S * K),
                    (double)blocked / (double)(S * K), (long long)bad_m, (long long)bad_d);
    }
}

void push3(std::vector<float> &v, double x, double y, double z)
{
    v.push_back((float)x);
    v.push_back((float)y);
    v.push_back((float)z);
}

// Height field on integer vertex coordinates: many rays lie in the planes
// x = const and y = const that hold vertex columns, so they meet triangle
// edges and vertices exactly (edge ties).
Scene terrain(uint64_t seed, int nqx, int nqy, int ns, int nk)
{
    Rng g(seed);
    Scene sc;
    sc.name = "terrain-ties";
    std::vector<double> h((nqx + 1) * (nqy + 1));
    for (auto &x : h)
        x = 2.0 * g.gauss();
    auto H = [&](int i, int j) { return h[j * (nqx + 1) + i]; };
    for (int j = 0; j < nqy; ++j)
        for (int i = 0; i < nqx; ++i) {
            const double p00[3] = {(double)i, (double)j, H(i, j)},
                         p10[3] = {(double)i + 1, (double)j, H(i + 1, j)},
                         p01[3] = {(double)i, (double)j + 1, H(i, j + 1)},
                         p11[3] = {(double)i + 1, (double)j + 1, H(i + 1, j + 1)};
            for (const double *p : {p00, p10, p11, p00, p11, p01})
                push3(sc.tris, p[0], p[1], p[2]);
        }
    // sources and sinks on a few vertex columns and rows: a third of the rays
    // lie in a plane x = i or y = j and cross the surface on a shared edge
    const int cols[4] = {nqx / 5, 2 * nqx / 5, 3 * nqx / 5, 4 * nqx / 5};
    const int rows[3] = {nqy / 4, nqy / 2, 3 * nqy / 4};
    auto top = [&](double x, double y) {
        const int i = std::min((int)x, nqx - 1), j = std::min((int)y, nqy - 1);
        return std::max(std::max(H(i, j), H(i + 1, j)), std::max(H(i, j + 1), H(i + 1, j + 1)));
    };
    for (int a = 0; a < ns; ++a) {
        const int i = a % 3 == 0 ? cols[(int)g.uni(0, 4)] : (int)g.uni(0, nqx);
        const int j = a % 3 == 1 ? rows[(int)g.uni(0, 3)] : (int)g.uni(0, nqy);
        push3(sc.src, i, j, top(i, j) + g.uni(0.5, 4.0));
    }
    for (int a = 0; a < nk; ++a) {
        double x = g.uni(0, nqx), y = g.uni(0, nqy);
        if (a % 3 == 0)
            x = cols[(int)g.uni(0, 4)];
        if (a % 3 == 1)
            y = rows[(int)g.uni(0, 3)];
        push3(sc.snk, x, y, top(x, y) - g.uni(-1.0, 2.0));
    }
    return sc;
}

// The terrain under a tall, mostly empty sky (a few small triangles float
// 30-40 m up, so the grid's upper blocks are empty): sources high above the
// surface, sinks near it, some on the vertex columns and rows -- rays cross
// long runs of empty macro blocks before they reach the surface, and some skim
// just above it.
Scene terrain_sky(uint64_t seed, int nqx, int nqy, int ns, int nk)
{
    Scene sc = terrain(seed, nqx, nqy, 0, 0);
    sc.name = "terrain-sky";
    Rng g(seed + 1000);
    for (int a = 0; a < 6; ++a) {
        const double x = g.uni(0, nqx), y = g.uni(0, nqy), z = g.uni(30, 40);
        push3(sc.tris, x, y, z);
        push3(sc.tris, x + 0.5, y, z);
        push3(sc.tris, x, y + 0.5, z + 0.2);
    }
    const int cols[3] = {nqx / 4, nqx / 2, 3 * nqx / 4};
    for (int a = 0; a < ns; ++a) {
        double x = g.uni(0, nqx), y = g.uni(0, nqy);
        if (a % 4 == 0)
            x = cols[(int)g.uni(0, 3)];
        push3(sc.src, x, y, g.uni(6.0, 36.0));
    }
    for (int a = 0; a < nk; ++a) {
        double x = g.uni(0, nqx), y = g.uni(0, nqy);
        if (a % 4 == 0)
            x = cols[(int)g.uni(0, 3)];
        push3(sc.snk, x, y, a % 2 ? g.uni(-4.0, 6.0) : g.uni(6.0, 36.0));
    }
    return sc;
}

// Points exactly on the grid's own cell boundaries (the float32 values the
// walk steps on, lo + cell * i): rays along cell planes, along cell edge lines
// and through cell corners, over the terrain of `base`.
Scene cell_boundaries(const Scene &base, float cells_per_tri, uint64_t seed, int ns, int nk)
{
    Scene sc;
    sc.name = "cell-bounds";
    sc.tris = base.tris;
    lem_grid_info gi;
    std::memset(&gi, 0, sizeof(gi));
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = __builtin_huge_valf();
        hi[a] = -__builtin_huge_valf();
    }
    for (size_t v = 0; v < sc.tris.size() / 3; ++v)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], sc.tris[3 * v + a]);
            hi[a] = std::max(hi[a], sc.tris[3 * v + a]);
        }
    grid_geometry(lo, hi, (int64_t)sc.tris.size() / 9, cells_per_tri, &gi);
    Rng g(seed);
    auto bnd = [&](int a, int i) { return gi.lo[a] + gi.cell[a] * (float)i; };
    auto pick = [&](int a) { return bnd(a, (int)g.uni(1, gi.n[a])); };
    const float zlo = hi[2] + 1.0f, zhi = lo[2] - 1.0f;
    for (int a = 0; a < ns; ++a)
        push3(sc.src, pick(0), pick(1), a % 2 ? zlo : pick(2));
    for (int a = 0; a < nk; ++a) {
        const int s = (int)g.uni(0, ns);
        float x = pick(0), y = pick(1), z = a % 2 ? zhi : pick(2);
        if (a % 3 == 0)
            x = sc.src[3 * s];      // the plane x = const of that source
        if (a % 5 == 0)
            y = sc.src[3 * s + 1];  // and, for some, the cell edge line
        push3(sc.snk, x, y, z);
    }
    return sc;
}

Scene clutter(uint64_t seed, int n_tris, int ns, int nk, double size)
{
    Rng g(seed);
    Scene sc;
    sc.name = "clutter";
    for (int t = 0; t < n_tris; ++t) {
        const double c[3] = {g.uni(-2, 2), g.uni(-2, 2), g.uni(-2, 2)};
        for (int v = 0; v < 3; ++v)
            push3(sc.tris, c[0] + g.uni(-size, size), c[1] + g.uni(-size, size),
                  c[2] + g.uni(-size, size));
    }
    for (int a = 0; a < ns; ++a)
        push3(sc.src, g.uni(-3, 3), g.uni(-3, 3), g.uni(-3, 3));
    for (int a = 0; a < nk; ++a)
        push3(sc.snk, g.uni(-3, 3), g.uni(-3, 3), g.uni(-3, 3));
    return sc;
}

// Rays along grid planes and axes: sources and sinks sharing a z (and some an
// x as well), so direction components are exactly zero.
Scene planes(uint64_t seed, int n_tris, int ns, int nk)
{
    Scene sc = clutter(seed, n_tris, 0, 0, 0.3);
    sc.name = "plane-rays";
    Rng g(seed + 1);
    for (int a = 0; a < ns; ++a)
        push3(sc.src, g.uni(-3, 3), g.uni(-3, 3), a % 2 ? 0.25 : g.uni(-3, 3));
    for (int a = 0; a < nk; ++a) {
        double x = g.uni(-3, 3), y = g.uni(-3, 3), z = a % 3 ? 0.25 : g.uni(-3, 3);
        if (a % 5 == 0)
            x = sc.src[3 * (a % ns)];
        push3(sc.snk, x, y, z);
    }
    // a few zero-length rays (a sink on a source)
    for (int a = 0; a < 3 && a < ns && a < nk; ++a)
        for (int c = 0; c < 3; ++c)
            sc.snk[3 * (nk - 1 - a) + c] = sc.src[3 * a + c];
    return sc;
}

// A 60 m scene with sinks 300 m to 30 km away (origins far outside the grid).
Scene far_scene(uint64_t seed, int n_tris, int ns, int nk)
{
    Rng g(seed);
    Scene sc;
    sc.name = "far-origins";
    for (int t = 0; t < n_tris; ++t) {
        const double c[3] = {g.uni(-30, 30), g.uni(-30, 30), g.uni(0, 20)};
        for (int v = 0; v < 3; ++v)
            push3(sc.tris, c[0] + g.uni(-3, 3), c[1] + g.uni(-3, 3), c[2] + g.uni(-3, 3));
    }
    for (int a = 0; a < ns; ++a)
        push3(sc.src, g.uni(-25, 25), g.uni(-25, 25), g.uni(1, 15));
    for (int a = 0; a < nk; ++a) {
        const double r = std::exp(g.uni(std::log(300.0), std::log(30000.0)));
        const double th = g.uni(0.0, 3.14159), ph = g.uni(-3.14159, 3.14159);
        push3(sc.snk, r * std::sin(th) * std::cos(ph), r * std::sin(th) * std::sin(ph),
              10.0 + r * std::cos(th));
    }
    return sc;
}

}  // namespace

int main(int argc, char **argv)
{
    // "quick": fewer rays (the sanitizer build)
    const bool quick = argc > 1 && std::strcmp(argv[1], "quick") == 0;
    const int q = quick ? 2 : 1;
    const Scene ter = terrain(11, 48, 24, 120 / q, 120 / q);
    const Scene cb1 = cell_boundaries(ter, 2.0f, 21, 90 / q, 90 / q);
    const Scene cb2 = cell_boundaries(ter, 1.0f, 22, 90 / q, 90 / q);
    const Scene clu = clutter(12, 1500, 70 / q, 70 / q, 0.3);
    const Scene pla = planes(13, 1200, 60 / q, 60 / q);
    const Scene far = far_scene(14, 800, 50 / q, 60 / q);
    const Scene sky = terrain_sky(15, 64, 48, 110 / q, 110 / q);
    for (float cpt : {1.0f, 2.0f, 4.0f})
        check_scene(ter, cpt, -1);
    check_scene(cb1, 2.0f, -1);
    check_scene(cb1, 2.0f, 0);
    check_scene(cb2, 1.0f, -1);
    for (int ms = 0; ms <= 4; ++ms)
        check_scene(clu, 2.0f, ms);
    check_scene(clu, 3.0f, -1);
    check_scene(pla, 2.0f, -1);
    check_scene(pla, 1.0f, 1);
    check_scene(far, 2.0f, -1);
    check_scene(far, 4.0f, 3);
    for (int ms = 0; ms <= 2; ++ms)
        check_scene(sky, 2.0f, ms);
    std::printf("total: %lld rays cast, %lld mismatches\n", (long long)g_rays,
                (long long)g_bad);
    return g_bad == 0 ? 0 : 1;
}
