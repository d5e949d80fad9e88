// lem_time.hip -- time-domain propagation of ray paths (gfx950).
//
// Replaces timedomainkernal (empropagation.py:1345-1612) behind TimeDomainv3
// (empropagation.py:2910-3170) and models/time_domain.py:
//   k_time_paths : per path row, the real polarisation vector carried along the
//                  path (projections and point weights, timedomainkernal's
//                  ray_component), scaled by |product of lossy_propagation
//                  factors|, and the arrival time (path length / c + the
//                  imaginary parts of the point weights, the reference's delay
//                  convention).  float64, with the float32 roundings the
//                  reference applies inside lossy_propagation.
//   k_time_wake  : wake time = min(1 s, earliest arrival) -- the reference's
//                  atomic minimum from an initial 1 s, taken over all paths
//                  before any path is placed (the reference reads it after a
//                  block-level barrier only, so its value is launch-order
//                  dependent).
//   k_time_map   : the excitation placed at each path's start sample
//                  floor((arrival - wake) / (1 / fs)) (Python's float floor
//                  division), summed per (source, sink) slot in path order --
//                  no atomics, deterministic.  out[slot][n][c] float64.
//   k_time_trace : the same summed over every path (the source and sink sums
//                  of time_domain.calculate_scattering), out[n][c].
#include "lem_device.h"
#include "lem_internal.h"
#include "../../include/lem.h"

#include <math.h>

#include <algorithm>

namespace lem {

constexpr int TB = 256;
constexpr double LIGHT_SPEED = 299792458.0;  // scipy.constants.c

// point i of the table, with the reference's negative-index wrap (a zero
// entry of the 1-based index reads point -1, the last one)
LEM_DEV int64_t wrap_point(int32_t one_based, int64_t N)
{
    int64_t i = (int64_t)one_based - 1;
    return i < 0 ? i + N : i;
}

// calc_sep (empropagation.py:82-93): float64 length of float32 differences
LEM_DEV double sep64(const float *__restrict__ pos, int64_t a, int64_t b)
{
    const double dx = (double)(pos[3 * b] - pos[3 * a]);
    const double dy = (double)(pos[3 * b + 1] - pos[3 * a + 1]);
    const double dz = (double)(pos[3 * b + 2] - pos[3 * a + 2]);
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// |lossy_propagation(a -> b)| (empropagation.py:768-801): length rounded to
// float32, float32 direction and obliquity dot product, 1/length rounded to
// float32 inside dG; magnitude of
//   -(1/2pi) (-(alpha + j beta) - 1/len) e^{-(alpha + j beta) len} / len (d.n_b).
LEM_DEV double loss_mag(const float *__restrict__ pos, const float *__restrict__ nrm, int64_t a,
                        int64_t b, double alpha, double beta)
{
    const float dx = pos[3 * b] - pos[3 * a], dy = pos[3 * b + 1] - pos[3 * a + 1],
                dz = pos[3 * b + 2] - pos[3 * a + 2];
    const double norm = sqrt((double)dx * dx + (double)dy * dy + (double)dz * dz);
    const float len = (float)norm;
    const float ux = (float)((double)dx / norm), uy = (float)((double)dy / norm),
                uz = (float)((double)dz / norm);
    float dot = ux * nrm[3 * b] + uy * nrm[3 * b + 1];
    dot = dot + uz * nrm[3 * b + 2];
    const double inv = 1.0 / (double)len;
    const double inv32 = (double)(float)inv;
    const double re = alpha + inv32, im = beta;
    return (1.0 / (2.0 * M_PI)) * sqrt(re * re + im * im) * exp(-alpha * (double)len) * inv *
           fabs((double)dot);
}

// (I - d d^T) v for the unit float64 direction a -> b (calc_dv into a float64
// array, then sourcelaunchtransformGPU's u/v projection)
LEM_DEV void project_dir(const float *__restrict__ pos, int64_t a, int64_t b, double v[3])
{
    const double dx = (double)(pos[3 * b] - pos[3 * a]);
    const double dy = (double)(pos[3 * b + 1] - pos[3 * a + 1]);
    const double dz = (double)(pos[3 * b + 2] - pos[3 * a + 2]);
    const double n = sqrt(dx * dx + dy * dy + dz * dz);
    const double ux = dx / n, uy = dy / n, uz = dz / n;
    const double p = v[0] * ux + v[1] * uy + v[2] * uz;
    v[0] -= p * ux;
    v[1] -= p * uy;
    v[2] -= p * uz;
}

// (I - n n^T) v with n the normalised point normal (the u/v basis of
// sourcelaunchtransformGPU is built from normalised cross products)
LEM_DEV void project_normal64(const float *__restrict__ nrm, int64_t a, double v[3])
{
    const double mx = nrm[3 * a], my = nrm[3 * a + 1], mz = nrm[3 * a + 2];
    const double m = sqrt(mx * mx + my * my + mz * mz);
    const double nx = mx / m, ny = my / m, nz = mz / m;
    const double p = v[0] * nx + v[1] * ny + v[2] * nz;
    v[0] -= p * nx;
    v[1] -= p * ny;
    v[2] -= p * nz;
}

__global__ __launch_bounds__(TB) void k_time_paths(
    const float *__restrict__ pos, const float *__restrict__ nrm, const float2 *__restrict__ w,
    int64_t N, const int32_t *__restrict__ index, int64_t V, int ncol, double alpha, double beta,
    double *__restrict__ amp, double *__restrict__ arrival, int32_t *__restrict__ err)
{
    const int64_t v = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (v >= V)
        return;
    const int32_t *row = index + v * ncol;
    bool ok = row[0] != 0;
    for (int i = 0; i < ncol; ++i)
        ok = ok && row[i] >= 0 && (int64_t)row[i] <= N;
    if (!ok) {
        err[0] = 1;
        const double nan = __longlong_as_double(0x7ff8000000000000ll);
        amp[3 * v] = amp[3 * v + 1] = amp[3 * v + 2] = arrival[v] = nan;
        return;
    }
    const int64_t p0 = wrap_point(row[0], N), p1 = wrap_point(row[1], N);
    double len = sep64(pos, p0, p1);
    double delay = (double)w[3 * p0].y;
    double rc[3] = {(double)w[3 * p0].x, (double)w[3 * p0 + 1].x, (double)w[3 * p0 + 2].x};
    if (row[1] != 0) {
        project_dir(pos, p0, p1, rc);
        rc[0] *= (double)w[3 * p1].x;
        rc[1] *= (double)w[3 * p1 + 1].x;
        rc[2] *= (double)w[3 * p1 + 2].x;
        delay += (double)w[3 * p1].y;
    }
    double mag = loss_mag(pos, nrm, p0, p1, alpha, beta);
    for (int i = 1; i < ncol - 1; ++i) {
        const int64_t a = wrap_point(row[i], N), b = wrap_point(row[i + 1], N);
        project_normal64(nrm, a, rc);
        len += sep64(pos, a, b);
        if (row[i + 1] != 0) {
            project_dir(pos, a, b, rc);
            rc[0] *= (double)w[3 * b].x;
            rc[1] *= (double)w[3 * b + 1].x;
            rc[2] *= (double)w[3 * b + 2].x;
            delay += (double)w[3 * b].y;
            mag *= loss_mag(pos, nrm, a, b, alpha, beta);
        }
    }
    amp[3 * v] = rc[0] * mag;
    amp[3 * v + 1] = rc[1] * mag;
    amp[3 * v + 2] = rc[2] * mag;
    arrival[v] = len / LIGHT_SPEED + delay;
}

// Order-preserving map of float64 to uint64 (for an atomic minimum).
LEM_DEV unsigned long long dkey(double d)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(d);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
LEM_DEV double dkey_inv(unsigned long long k)
{
    const unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)u);
}

__global__ __launch_bounds__(TB) void k_time_wake(const double *__restrict__ arrival, int64_t V,
                                                  unsigned long long *__restrict__ key)
{
    __shared__ unsigned long long s[TB];
    unsigned long long m = dkey(1.0);
    for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < V; i += (int64_t)gridDim.x * TB) {
        const double a = arrival[i];
        if (a == a) {  // NaN arrivals take no part
            const unsigned long long k = dkey(a);
            m = k < m ? k : m;
        }
    }
    s[threadIdx.x] = m;
    __syncthreads();
    for (int off = TB / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            s[threadIdx.x] = s[threadIdx.x + off] < s[threadIdx.x] ? s[threadIdx.x + off]
                                                                   : s[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicMin(key, s[0]);
}

__global__ void k_time_wake_init(unsigned long long *__restrict__ key)
{
    if (threadIdx.x == 0)
        *key = dkey(1.0);
}

__global__ void k_time_wake_finish(unsigned long long *__restrict__ key)
{
    if (threadIdx.x == 0)
        *reinterpret_cast<double *>(key) = dkey_inv(*key);
}

// Python's float floor division a // b (CPython float_floor_div, which numba
// follows): fmod-based, exact for the integer-valued results used as indices.
LEM_DEV double py_floordiv(double a, double b)
{
    double mod = fmod(a, b);
    double div = (a - mod) / b;
    if (mod != 0.0) {
        if ((b < 0.0) != (mod < 0.0))
            div -= 1.0;
    }
    if (div != 0.0) {
        double fl = floor(div);
        if (div - fl > 0.5)
            fl += 1.0;
        return fl;
    }
    return copysign(0.0, a / b);
}

// start sample of each path
__global__ __launch_bounds__(TB) void k_time_bins(const double *__restrict__ arrival, int64_t V,
                                                  const double *__restrict__ wake, double sep,
                                                  int64_t *__restrict__ bin)
{
    const int64_t v = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (v >= V)
        return;
    const double b = py_floordiv(arrival[v] - *wake, sep);
    bin[v] = b < 9.0e18 ? (int64_t)b : (int64_t)9e18;
}

// One thread per (segment, sample n): sum over the segment's paths, in path
// order, of excitation[n - bin] * amp.  Segments are the (source, sink) slots
// (paths sorted stably by slot).  Grid: (ceil(N / TB), nseg).
__global__ __launch_bounds__(TB) void k_time_map(
    const double *__restrict__ amp, const int64_t *__restrict__ bin,
    const int64_t *__restrict__ order, const int64_t *__restrict__ seg_start,
    const int64_t *__restrict__ seg_slot, const double *__restrict__ exc, int64_t E, int64_t N,
    double *__restrict__ out)
{
    const int64_t n = (int64_t)blockIdx.x * TB + threadIdx.x;
    const int64_t g = blockIdx.y;
    if (n >= N)
        return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int64_t q = seg_start[g]; q < seg_start[g + 1]; ++q) {
        const int64_t p = order[q];
        const int64_t j = n - bin[p];
        if (j >= 0 && j < E) {
            const double e = exc[j];
            a0 += e * amp[3 * p];
            a1 += e * amp[3 * p + 1];
            a2 += e * amp[3 * p + 2];
        }
    }
    double *o = out + (seg_slot[g] * N + n) * 3;
    o[0] = a0;
    o[1] = a1;
    o[2] = a2;
}

// Whole trace: one thread per sample, every path in index order, paths staged
// through LDS 256 at a time.
__global__ __launch_bounds__(TB) void k_time_trace(const double *__restrict__ amp,
                                                   const int64_t *__restrict__ bin, int64_t V,
                                                   const double *__restrict__ exc, int64_t E,
                                                   int64_t N, double *__restrict__ out)
{
    __shared__ double sa[TB][3];
    __shared__ int64_t sb[TB];
    const int64_t n = (int64_t)blockIdx.x * TB + threadIdx.x;
    const int64_t n0 = (int64_t)blockIdx.x * TB;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int64_t p0 = 0; p0 < V; p0 += TB) {
        __syncthreads();
        const int64_t p = p0 + threadIdx.x;
        if (p < V) {
            sb[threadIdx.x] = bin[p];
            sa[threadIdx.x][0] = amp[3 * p];
            sa[threadIdx.x][1] = amp[3 * p + 1];
            sa[threadIdx.x][2] = amp[3 * p + 2];
        } else {
            sb[threadIdx.x] = -(int64_t)1 << 62;  // never in range
        }
        __syncthreads();
        const int cnt = (int)(V - p0 < TB ? V - p0 : TB);
        for (int q = 0; q < cnt; ++q) {
            const int64_t b = sb[q];
            if (b > n0 + TB - 1 || b + E <= n0)
                continue;  // the whole block is outside this path's window
            const int64_t j = n - b;
            if (j >= 0 && j < E) {
                const double e = exc[j];
                a0 += e * sa[q][0];
                a1 += e * sa[q][1];
                a2 += e * sa[q][2];
            }
        }
    }
    if (n < N) {
        out[3 * n] = a0;
        out[3 * n + 1] = a1;
        out[3 * n + 2] = a2;
    }
}

}  // namespace lem

using namespace lem;

extern "C" int lem_time_paths(const float *pos, const float *nrm, const float *w, int64_t N,
                              const int32_t *index, int64_t V, int ncol, double alpha, double beta,
                              double *amp, double *arrival, int32_t *err, void *stream)
{
    clear_error();
    LEM_REQUIRE(ncol >= 2 && ncol <= 8, "lem_time_paths: index needs 2 to 8 columns");
    LEM_REQUIRE(V >= 0 && N >= 0, "lem_time_paths: negative size");
    if (V == 0)
        return 0;
    LEM_REQUIRE(N > 0 && pos && nrm && w && index && amp && arrival && err,
                "lem_time_paths: null pointer");
    hipLaunchKernelGGL(k_time_paths, dim3((unsigned)ceil_div(V, TB)), dim3(TB), 0,
                       as_stream(stream), pos, nrm, reinterpret_cast<const float2 *>(w), N, index,
                       V, ncol, alpha, beta, amp, arrival, err);
    return check_launch("k_time_paths");
}

extern "C" int lem_time_wake(const double *arrival, int64_t V, double *wake, void *stream)
{
    clear_error();
    LEM_REQUIRE(V >= 0 && wake && (V == 0 || arrival), "lem_time_wake: bad arguments");
    hipStream_t st = as_stream(stream);
    unsigned long long *key = reinterpret_cast<unsigned long long *>(wake);
    hipLaunchKernelGGL(k_time_wake_init, dim3(1), dim3(64), 0, st, key);
    if (V > 0) {
        const int64_t nb = std::min<int64_t>(ceil_div(V, TB), 1024);
        hipLaunchKernelGGL(k_time_wake, dim3((unsigned)nb), dim3(TB), 0, st, arrival, V, key);
    }
    hipLaunchKernelGGL(k_time_wake_finish, dim3(1), dim3(64), 0, st, key);
    return check_launch("k_time_wake");
}

extern "C" int lem_time_bins(const double *arrival, int64_t V, const double *wake,
                             double sampling_freq, int64_t *bin, void *stream)
{
    clear_error();
    LEM_REQUIRE(V >= 0 && sampling_freq > 0.0, "lem_time_bins: bad arguments");
    if (V == 0)
        return 0;
    LEM_REQUIRE(arrival && wake && bin, "lem_time_bins: null pointer");
    hipLaunchKernelGGL(k_time_bins, dim3((unsigned)ceil_div(V, TB)), dim3(TB), 0,
                       as_stream(stream), arrival, V, wake, 1.0 / sampling_freq, bin);
    return check_launch("k_time_bins");
}

extern "C" int lem_time_map(const double *amp, const int64_t *bin, const int64_t *order,
                            const int64_t *seg_start, const int64_t *seg_slot, int64_t nseg,
                            const double *excitation, int64_t E, int64_t N, double *out,
                            void *stream)
{
    clear_error();
    LEM_REQUIRE(nseg >= 0 && E >= 0 && N >= 0, "lem_time_map: negative size");
    if (nseg == 0 || N == 0)
        return 0;
    LEM_REQUIRE(amp && bin && order && seg_start && seg_slot && out && (E == 0 || excitation),
                "lem_time_map: null pointer");
    // at most 65535 segments per launch (grid y)
    for (int64_t g0 = 0; g0 < nseg; g0 += 65535) {
        const int64_t ng = std::min<int64_t>(nseg - g0, 65535);
        dim3 grid((unsigned)ceil_div(N, TB), (unsigned)ng);
        hipLaunchKernelGGL(k_time_map, grid, dim3(TB), 0, as_stream(stream), amp, bin, order,
                           seg_start + g0, seg_slot + g0, excitation, E, N, out);
        const int rc = check_launch("k_time_map");
        if (rc)
            return rc;
    }
    return 0;
}

extern "C" int lem_time_trace(const double *amp, const int64_t *bin, int64_t V,
                              const double *excitation, int64_t E, int64_t N, double *out,
                              void *stream)
{
    clear_error();
    LEM_REQUIRE(V >= 0 && E >= 0 && N >= 0, "lem_time_trace: negative size");
    if (N == 0)
        return 0;
    LEM_REQUIRE(out && (V == 0 || (amp && bin)) && (E == 0 || excitation),
                "lem_time_trace: null pointer");
    hipLaunchKernelGGL(k_time_trace, dim3((unsigned)ceil_div(N, TB)), dim3(TB), 0,
                       as_stream(stream), amp, bin, V, excitation, E, N, out);
    return check_launch("k_time_trace");
}
