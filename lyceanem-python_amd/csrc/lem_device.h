// lem_device.h -- device-side building blocks shared by the gfx950 kernels.
//
// Everything here is compiled with -ffp-contract=off and correctly rounded
// float32 division / square root, so each float32 expression below is
// evaluated exactly as written.  That is what makes the hit masks
// reproducible bit for bit against the float32 reference arithmetic.
//
// The same code also compiles as host C++ for the CPU build of the grid walk
// (LEM_HOST_BUILD): that build supplies LEM_DEV, the HIP vector types and the
// three hardware estimates below.
#pragma once

#ifndef LEM_HOST_BUILD
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#ifndef LEM_DEV
#define LEM_DEV __device__ __forceinline__
#endif

namespace lem {

#ifndef LEM_HOST_BUILD
// Hardware estimates (v_rcp_f64, v_rcp_f32, v_sqrt_f32); every use refines them
// or tolerates their ~1 ulp error.
LEM_DEV double rcp_estimate(double a) { return __builtin_amdgcn_rcp(a); }
LEM_DEV float rcpf_estimate(float a) { return __builtin_amdgcn_rcpf(a); }
LEM_DEV float sqrtf_estimate(float a) { return __builtin_amdgcn_sqrtf(a); }
#endif

// ---------------------------------------------------------------------------
// Ray charging (calc_dv_norm, math_functions.py:148-156; points_to_rays,
// raycasting.cuh:80-95): float32 difference, left-to-right sum of squares,
// correctly rounded sqrt and divisions.
// ---------------------------------------------------------------------------
struct Ray {
    float ox, oy, oz, dx, dy, dz, len;
};

LEM_DEV Ray charge(float ox, float oy, float oz, float tx, float ty, float tz)
{
    Ray r;
    float dx = tx - ox, dy = ty - oy, dz = tz - oz;
    float s = dx * dx + dy * dy;
    s = s + dz * dz;
    float L = sqrtf(s);
    r.ox = ox;
    r.oy = oy;
    r.oz = oz;
    r.dx = dx / L;
    r.dy = dy / L;
    r.dz = dz / L;
    r.len = L;
    return r;
}

// Triangle in the form the predicates consume: v0 and the two edges, which the
// reference forms with the same float32 subtractions inside `hit`
// (rayfunctions.py:44-50, raycasting_brute_force.cuh:38-40).
struct Tri {
    float v0x, v0y, v0z, e1x, e1y, e1z, e2x, e2y, e2z;
};

LEM_DEV Tri make_tri(const float *t9)
{
    Tri t;
    t.v0x = t9[0];
    t.v0y = t9[1];
    t.v0z = t9[2];
    t.e1x = t9[3] - t9[0];
    t.e1y = t9[4] - t9[1];
    t.e1z = t9[5] - t9[2];
    t.e2x = t9[6] - t9[0];
    t.e2y = t9[7] - t9[1];
    t.e2z = t9[8] - t9[2];
    return t;
}

// Largest float strictly below the double 1e-7 (float(1e-7) rounds up).
// `-1e-7 < A < 1e-7` in float64 is exactly `|A| <= NUMBA_EPS_BELOW` in float32.
constexpr float NUMBA_EPS_BELOW = 9.999999406e-08f;

// float64 reciprocal of a float32 value, equal bit for bit to the IEEE
// quotient 1.0 / (double)a for every finite non-zero float32 a (and NaN for
// NaN): the hardware estimate refined by two Newton steps, y += y * (1 - a y)
// with fused operations.  A float32 a has a 24-bit significand, so 1/a is
// never a tie between two doubles; an exhaustive GPU test compares the two
// over every float32 bit pattern.  About a third of the instructions of the
// general float64 division (no scaling steps: 1/a stays far inside the double
// range for any float32 a).  For a = +-inf the quotient is +-0 and this gives
// NaN; hit_numba cannot tell them apart (see there).
LEM_DEV double recip_f32_to_f64(float a)
{
    const double ad = (double)a;
    double y = rcp_estimate(ad);
    y = __builtin_fma(y, __builtin_fma(-ad, y, 1.0), y);
    return __builtin_fma(y, __builtin_fma(-ad, y, 1.0), y);
}

// ---------------------------------------------------------------------------
// numba `hit` (rayfunctions.py:37-104): two-sided Moller-Trumbore with
// float32 edges / cross / dot and float64 F = 1/A, U, V, U+V and t.
//
// The float64 tests are decided in float32 wherever that is provably exact:
//   U < 0      <=>  du*sign(A) < 0
//   U > 1      <=>  du*sign(A) > |A|   (du, A float32: du/A != 1 differs from 1
//                                       by >= 2^-24, far above float64 rounding,
//                                       and du == A gives F*du <= 1 in float64)
//   V < 0      <=>  dv*sign(A) < 0
//   U + V > 1  decided from the float32 residual du'+dv'-|A| outside a band of
//              1e-6 relative; inside the band the float64 expression is evaluated
//              exactly as the reference writes it.
// t is always evaluated in float64 as F*dt (only triangles whose plane crossing
// lies inside the triangle get this far).
// Returns true for an accepted hit in the window (2e-7, dist - 2e-7).
// ---------------------------------------------------------------------------
LEM_DEV bool hit_numba(const Ray &r, float dist, const Tri &T, double &t_out)
{
    // The float32 part is evaluated in full and its three early outs (A, U, V)
    // are combined into one test: the values and decisions are those of the
    // sequential form, without a divergent branch per test.
    float px = r.dy * T.e2z - r.dz * T.e2y;
    float py = r.dz * T.e2x - r.dx * T.e2z;
    float pz = r.dx * T.e2y - r.dy * T.e2x;
    float A = px * T.e1x + py * T.e1y;
    A = A + pz * T.e1z;
    float Aa = fabsf(A);
    float tx = r.ox - T.v0x, ty = r.oy - T.v0y, tz = r.oz - T.v0z;
    float du = tx * px + ty * py;
    du = du + tz * pz;
    float sg = __builtin_copysignf(1.0f, A);
    float dus = du * sg;
    float qx = ty * T.e1z - tz * T.e1y;
    float qy = tz * T.e1x - tx * T.e1z;
    float qz = tx * T.e1y - ty * T.e1x;
    float dv = r.dx * qx + r.dy * qy;
    dv = dv + r.dz * qz;
    float dvs = dv * sg;
    // negations of the sequential tests (a NaN passes them, as it does there)
    const bool cand = !(Aa <= NUMBA_EPS_BELOW) & !(dus < 0.0f) & !(dus > Aa) & !(dvs < 0.0f);
    if (!cand)
        return false;
    float s1 = dus + dvs;
    float s2 = s1 - Aa;
    // == 1.0 / (double)A for finite A.  An infinite A (float32 overflow of the
    // products) makes F NaN instead of 0: U + V > 1 is then false for both
    // (NaN, or 0 / NaN), and the window test on t = F dt fails for both.
    double F = recip_f32_to_f64(A);
    if (fabsf(s2) > 1e-6f * (s1 + Aa)) {
        if (s2 > 0.0f)
            return false;
    } else {
        double U = F * (double)du;
        double V = F * (double)dv;
        if ((U + V) > 1.0)
            return false;
    }
    float dt = T.e2x * qx + T.e2y * qy;
    dt = dt + T.e2z * qz;
    double t = F * (double)dt;
    if (t > 2e-7 && t < (double)dist - 2e-7) {
        t_out = t;
        return true;
    }
    return false;
}

// ---------------------------------------------------------------------------
// Native CUDA `hit` (raycasting.cuh:19-68): back-face culled, eps = 1e-6f,
// all float32, F = (float)(1.0/A) which equals the correctly rounded float32
// reciprocal (the float64 quotient of a float32 can never sit within 2^-53 of
// a float32 rounding boundary).
// ---------------------------------------------------------------------------
LEM_DEV bool hit_native(const Ray &r, const Tri &T)
{
    float px = r.dy * T.e2z - r.dz * T.e2y;
    float py = r.dz * T.e2x - r.dx * T.e2z;
    float pz = r.dx * T.e2y - r.dy * T.e2x;
    float A = T.e1x * px + T.e1y * py;
    A = A + T.e1z * pz;
    if (fabsf(A) < 1e-6f)
        return false;
    if (A < 0.0f)
        return false;
    float tx = r.ox - T.v0x, ty = r.oy - T.v0y, tz = r.oz - T.v0z;
    float du = tx * px + ty * py;
    du = du + tz * pz;
    float F = 1.0f / A;
    float U = F * du;
    if (U < 0.0f || U > 1.0f)
        return false;
    float qx = ty * T.e1z - tz * T.e1y;
    float qy = tz * T.e1x - tx * T.e1z;
    float qz = tx * T.e1y - ty * T.e1x;
    float dv = r.dx * qx + r.dy * qy;
    dv = dv + r.dz * qz;
    float V = F * dv;
    float UV = U + V;
    if (V < 0.0f || UV > 1.0f)
        return false;
    float dt = T.e2x * qx + T.e2y * qy;
    dt = dt + T.e2z * qz;
    float t = F * dt;
    float hi = r.len - 2e-6f;
    return t > 2e-6f && t < hi;
}

// ---------------------------------------------------------------------------
// Field math.
// Path length and phase in float64 (the phase of a 1 km path is ~6e4 rad, so
// float32 would lose ~4e-3 rad); everything else float32.
// ---------------------------------------------------------------------------
struct C3 {
    float xr, xi, yr, yi, zr, zi;
};

// e^{-j beta r} and 1/r for a path of float64 length components.
// r = |delta| by one float64 Newton step on the float32 square root; 1/r is the
// float32 reciprocal of that square root.  The phase is carried in quarter
// turns, t = -beta r (2/pi): adding and removing 1.5 * 2^52 rounds t to the
// nearest integer q (exact for |t| < 2^51) and leaves q mod 4 in the low
// mantissa bits; the remainder x = t - q (|x| <= 1/2 + 2^-52) is the fused
// r (-beta 2/pi) - q, rounded once.
// Sine / cosine of (pi/2) x come from their Taylor series to x^9 / x^10
// (truncation error < 2e-9 on |x| <= pi/4), then the quadrant swap.
struct Phase {
    float c, s;     // cos(-beta r), sin(-beta r)
    float inv_r;    // 1/r (float32)
    double r;       // float64 path length
};

LEM_DEV void sincos_reduced(float x, float &s, float &c)
{
    float x2 = x * x;
    float sp = __builtin_fmaf(x2, 2.75573192e-6f, -1.98412698e-4f);
    sp = __builtin_fmaf(x2, sp, 8.33333333e-3f);
    sp = __builtin_fmaf(x2, sp, -1.66666667e-1f);
    s = __builtin_fmaf(x * x2, sp, x);
    float cp = __builtin_fmaf(x2, -2.75573192e-7f, 2.48015873e-5f);
    cp = __builtin_fmaf(x2, cp, -1.38888889e-3f);
    cp = __builtin_fmaf(x2, cp, 4.16666667e-2f);
    cp = __builtin_fmaf(x2, cp, -0.5f);
    c = __builtin_fmaf(x2, cp, 1.0f);
}

LEM_DEV Phase phase_of(double dx, double dy, double dz, double beta)
{
    Phase p;
    double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
    float r0 = sqrtf_estimate((float)r2);   // ~1 ulp seed
    float ir0 = rcpf_estimate(r0);
    double r0d = (double)r0;
    double e = __builtin_fma(-r0d, r0d, r2);
    double r = __builtin_fma(e, (double)(0.5f * ir0), r0d);
    p.r = r;
    p.inv_r = ir0;
    const double NEG_TWO_OVER_PI = -0.63661977236758134308;
    const double ROUND = 6755399441055744.0;  // 1.5 * 2^52
    const double b = beta * NEG_TWO_OVER_PI;  // quarter turns per unit length
    // t = r b; q = rint(t) from the fused product, remainder t - q rounded once
    double m = __builtin_fma(r, b, ROUND);
    double q = m - ROUND;
    float x = (float)__builtin_fma(r, b, -q) * 1.57079632679489662f;
    float s, c;
    sincos_reduced(x, s, c);
    // q mod 4 (two's complement of q in the low word of m)
    const int qi = (int)(unsigned)__builtin_bit_cast(unsigned long long, m) & 3;
    // (cos, sin) of x + qi pi/2: quadrants 0..3 give (c, s), (-s, c), (-c, -s),
    // (s, -c) -- selected, not branched on
    const float cs = (qi & 1) ? s : c, sn = (qi & 1) ? c : s;
    p.c = ((qi + 1) & 2) ? -cs : cs;
    p.s = (qi & 2) ? -sn : sn;
    return p;
}

// Scalar leg factor of lossy_propagation (empropagation.py:768-801) /
// em_wave (em.cuh:64-120):
//   loss = -(1/2pi) * (-(alpha + j beta) - 1/r) * e^{-(alpha + j beta) r}/r * cos
// with cos = d . n_end (raw end normal, no clipping).  Returned as (re, im).
// LOSSY = false: alpha is 0 (no attenuation term is evaluated).  The field
// arithmetic may contract a * b + c into one fused operation.
template <bool LOSSY = true>
LEM_DEV void leg_factor(const Phase &p, float alpha, float beta_f, float cosang,
                        float &lr, float &li)
{
#pragma clang fp contract(fast)
    const float FRONT = -0.15915494309189535f;  // -(1/(2 pi))
    // FRONT cos e^{-alpha r}/r e^{-j beta r}
    float amp = p.inv_r * (FRONT * cosang);
    if (LOSSY)
        amp *= expf(-alpha * (float)p.r);
    const float Gr = amp * p.c, Gi = amp * p.s;
    // times (-alpha - 1/r) - j beta
    const float ar = LOSSY ? -alpha - p.inv_r : -p.inv_r;
    lr = ar * Gr + beta_f * Gi;
    li = ar * Gi - beta_f * Gr;
}

// (I - d d^T) w for a complex 3-vector: the closed form of the u/v launch
// basis of sourcelaunchtransformGPU (empropagation.py:178-249) and ray_launch
// (em.cuh:18-61), (E.u)u + (E.v)v with u, v, d orthonormal.
LEM_DEV C3 project(const C3 &w, float dx, float dy, float dz)
{
#pragma clang fp contract(fast)
    float pr = w.xr * dx + w.yr * dy + w.zr * dz;
    float pi = w.xi * dx + w.yi * dy + w.zi * dz;
    C3 o;
    o.xr = w.xr - dx * pr;
    o.xi = w.xi - dx * pi;
    o.yr = w.yr - dy * pr;
    o.yi = w.yi - dy * pi;
    o.zr = w.zr - dz * pr;
    o.zi = w.zi - dz * pi;
    return o;
}

// Kahan-compensated float32 accumulator for one complex 3-vector.  The source
// and sink loops add plain float32 partials of SUM_GROUP consecutive terms and
// fold each partial into the accumulator (a group's rounding is at most
// (SUM_GROUP - 1) ulp of its terms' magnitudes; the compensated sum keeps the
// long loops from growing it).
constexpr int SUM_GROUP = 16;

struct KahanC3 {
    float s[6];
    float c[6];
    LEM_DEV void zero()
    {
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            s[i] = 0.0f;
            c[i] = 0.0f;
        }
    }
    LEM_DEV void add(const float v[6])
    {
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            float y = v[i] - c[i];
            float t = s[i] + y;
            c[i] = (t - s[i]) - y;
            s[i] = t;
        }
    }
};

}  // namespace lem
