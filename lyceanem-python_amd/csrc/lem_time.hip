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
//   trace        : the same summed over every path (the source and sink sums
//                  of time_domain.calculate_scattering), out[n][c]: a stable
//                  radix sort of the paths by start sample, per-sample sums of
//                  the amplitudes in path order, then the excitation
//                  convolution -- O(V + N E) instead of the reference's V x E
//                  atomic additions, and deterministic.
#include "lem_device.h"
#include "lem_internal.h"
#include "../../include/lem.h"

#include <math.h>

#include <algorithm>

#include <hipcub/hipcub.hpp>

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
    // lossless: e^{-0 len} = 1 without evaluating the exponential
    const double att = alpha != 0.0 ? exp(-alpha * (double)len) : 1.0;
    return (1.0 / (2.0 * M_PI)) * sqrt(re * re + im * im) * att * inv * fabs((double)dot);
}

// (I - d d^T) v for the unit float64 direction a -> b (calc_dv into a float64
// array, then sourcelaunchtransformGPU's u/v projection)
LEM_DEV void project_dir(const float *__restrict__ pos, int64_t a, int64_t b, double v[3])
{
    const double dx = (double)(pos[3 * b] - pos[3 * a]);
    const double dy = (double)(pos[3 * b + 1] - pos[3 * a + 1]);
    const double dz = (double)(pos[3 * b + 2] - pos[3 * a + 2]);
    const double in = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
    const double ux = dx * in, uy = dy * in, uz = dz * in;
    const double p = v[0] * ux + v[1] * uy + v[2] * uz;
    v[0] -= p * ux;
    v[1] -= p * uy;
    v[2] -= p * uz;
}

// A magnetic first point (timedomainkernal, empropagation.py:1395-1419):
// v = Re(w) x d / eta, d the float64 unit direction a -> b
LEM_DEV void magnetic_launch(const float *__restrict__ pos, int64_t a, int64_t b, double eta,
                             double v[3])
{
    const double dx = (double)(pos[3 * b] - pos[3 * a]);
    const double dy = (double)(pos[3 * b + 1] - pos[3 * a + 1]);
    const double dz = (double)(pos[3 * b + 2] - pos[3 * a + 2]);
    const double n = sqrt(dx * dx + dy * dy + dz * dz);
    const double ux = dx / n, uy = dy / n, uz = dz / n;
    const double mx = (v[1] * uz - v[2] * uy) / eta;
    const double my = (v[2] * ux - v[0] * uz) / eta;
    const double mz = (v[0] * uy - v[1] * ux) / eta;
    v[0] = mx;
    v[1] = my;
    v[2] = mz;
}

// (I - n n^T) v with n the normalised point normal (the u/v basis of
// sourcelaunchtransformGPU is built from normalised cross products)
LEM_DEV void project_normal64(const float *__restrict__ nrm, int64_t a, double v[3])
{
    const double mx = nrm[3 * a], my = nrm[3 * a + 1], mz = nrm[3 * a + 2];
    const double im = 1.0 / sqrt(mx * mx + my * my + mz * mz);
    const double nx = mx * im, ny = my * im, nz = mz * im;
    const double p = v[0] * nx + v[1] * ny + v[2] * nz;
    v[0] -= p * nx;
    v[1] -= p * ny;
    v[2] -= p * nz;
}

__global__ __launch_bounds__(TB) void k_time_paths(
    const float *__restrict__ pos, const float *__restrict__ nrm, const float2 *__restrict__ w,
    const float *__restrict__ eta, int64_t N, const int32_t *__restrict__ index, int64_t V,
    int ncol, double alpha, double beta,
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
    if (eta && eta[p0] > 0.0f)
        magnetic_launch(pos, p0, p1, (double)eta[p0], rc);
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

// Trace: sort keys (start sample clamped to [0, N]; N = never placed).
__global__ __launch_bounds__(TB) void k_trace_keys(const int64_t *__restrict__ bin, int64_t V,
                                                   int64_t N, int32_t *__restrict__ key,
                                                   int32_t *__restrict__ val)
{
    const int64_t p = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (p >= V)
        return;
    // keys 0 .. N - 1 (a 2^16-sample trace sorts on 16 bits); a path that
    // starts at or past the trace's end keeps key N - 1 and is marked -1
    const int64_t b = bin[p];
    key[p] = (int32_t)(b < 0 ? 0 : (b >= N ? N - 1 : b));
    val[p] = b >= N ? -1 : (int32_t)p;
}

// start[b] = first sorted position with key >= b, for b in [0, N]
__global__ __launch_bounds__(TB) void k_trace_starts(const int32_t *__restrict__ key, int64_t V,
                                                     int64_t N, int32_t *__restrict__ start)
{
    const int64_t b = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (b > N)
        return;
    int64_t lo = 0, hi = V;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (key[mid] < b)
            lo = mid + 1;
        else
            hi = mid;
    }
    start[b] = (int32_t)lo;
}

// H[b] = sum of the amplitudes of the paths starting at sample b.  One
// wavefront per sample: lane l sums the sample's paths l, l + 64, ... in path
// order, then the 64 partial sums meet in a fixed butterfly -- the order is
// fixed, so the result is deterministic.
__global__ __launch_bounds__(TB) void k_trace_hist(const double *__restrict__ amp,
                                                   const int32_t *__restrict__ val,
                                                   const int32_t *__restrict__ start, int64_t N,
                                                   double *__restrict__ H)
{
    const int64_t b = (int64_t)blockIdx.x * (TB / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (b >= N)
        return;  // wave-uniform
    double h0 = 0.0, h1 = 0.0, h2 = 0.0;
    const int32_t e = start[b + 1];
    for (int32_t i = start[b] + lane; i < e; i += 64) {
        const int64_t p = val[i];
        if (p < 0)
            continue;  // past the trace's end
        h0 += amp[3 * p];
        h1 += amp[3 * p + 1];
        h2 += amp[3 * p + 2];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        h0 += __shfl_xor(h0, off);
        h1 += __shfl_xor(h1, off);
        h2 += __shfl_xor(h2, off);
    }
    if (lane == 0) {
        H[3 * b] = h0;
        H[3 * b + 1] = h1;
        H[3 * b + 2] = h2;
    }
}

// out[n] = sum_{j <= min(E - 1, n)} exc[j] H[n - j]
__global__ __launch_bounds__(TB) void k_trace_conv(const double *__restrict__ H,
                                                   const double *__restrict__ exc, int64_t E,
                                                   int64_t N, double *__restrict__ out)
{
    const int64_t n = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (n >= N)
        return;
    const int64_t J = n + 1 < E ? n + 1 : E;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int64_t j = 0; j < J; ++j) {
        const double e = exc[j];
        const double *h = H + 3 * (n - j);
        a0 += e * h[0];
        a1 += e * h[1];
        a2 += e * h[2];
    }
    out[3 * n] = a0;
    out[3 * n + 1] = a1;
    out[3 * n + 2] = a2;
}

struct TraceWs {
    int32_t *k0, *k1, *v0, *v1, *start;
    double *H;
    void *cub;
    size_t cub_bytes;
};

inline int key_bits(int64_t N)
{
    int b = 1;
    while (b < 31 && ((int64_t)1 << b) <= N)
        ++b;
    return b;
}

inline int64_t align256(int64_t x) { return (x + 255) / 256 * 256; }

// Lays the trace workspace out from base (or only sizes it when base is null).
inline int64_t trace_layout(int64_t V, int64_t N, size_t cub_bytes, char *base, TraceWs *L)
{
    int64_t off = 0;
    auto take = [&](int64_t bytes) {
        char *p = base ? base + off : nullptr;
        off += align256(bytes);
        return p;
    };
    char *k0 = take(4 * V), *k1 = take(4 * V), *v0 = take(4 * V), *v1 = take(4 * V);
    char *st = take(4 * (N + 2));
    char *H = take(8 * 3 * N);
    char *cub = take((int64_t)cub_bytes);
    if (L) {
        L->k0 = reinterpret_cast<int32_t *>(k0);
        L->k1 = reinterpret_cast<int32_t *>(k1);
        L->v0 = reinterpret_cast<int32_t *>(v0);
        L->v1 = reinterpret_cast<int32_t *>(v1);
        L->start = reinterpret_cast<int32_t *>(st);
        L->H = reinterpret_cast<double *>(H);
        L->cub = cub;
        L->cub_bytes = cub_bytes;
    }
    return off;
}

// Scratch bytes of the radix sort, or -1 when the size query fails.
inline int64_t cub_sort_bytes(int64_t V, int64_t N)
{
    size_t bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const int32_t *)nullptr,
                                           (int32_t *)nullptr, (const int32_t *)nullptr,
                                           (int32_t *)nullptr, (int)V, 0,
                                           key_bits(N - 1)) != hipSuccess)
        return -1;
    return (int64_t)bytes;
}

}  // namespace lem

using namespace lem;

extern "C" int lem_time_paths(const float *pos, const float *nrm, const float *w,
                              const float *eta, int64_t N,
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
                       as_stream(stream), pos, nrm, reinterpret_cast<const float2 *>(w), eta, N,
                       index, V, ncol, alpha, beta, amp, arrival, err);
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

extern "C" int64_t lem_time_trace_workspace(int64_t V, int64_t N)
{
    if (V < 0 || N < 0 || V > 2000000000LL || N > 2000000000LL)
        return -1;
    const int64_t cub = cub_sort_bytes(V > 0 ? V : 1, N);
    if (cub < 0)
        return -1;
    return trace_layout(V, N, (size_t)cub, nullptr, nullptr);
}

extern "C" int lem_time_trace(const double *amp, const int64_t *bin, int64_t V,
                              const double *excitation, int64_t E, int64_t N, double *out,
                              void *workspace, int64_t workspace_bytes, void *stream)
{
    clear_error();
    LEM_REQUIRE(V >= 0 && E >= 0 && N >= 0, "lem_time_trace: negative size");
    LEM_REQUIRE(V <= 2000000000LL && N <= 2000000000LL, "lem_time_trace: too many paths or samples");
    if (N == 0)
        return 0;
    LEM_REQUIRE(out && (V == 0 || (amp && bin)) && (E == 0 || excitation),
                "lem_time_trace: null pointer");
    const int64_t Vs = V > 0 ? V : 1;
    const int64_t cub = cub_sort_bytes(Vs, N);
    if (cub < 0) {
        set_error("lem_time_trace: radix sort size query failed");
        return E_HIP;
    }
    const size_t cub_bytes = (size_t)cub;
    const int64_t need = trace_layout(Vs, N, cub_bytes, nullptr, nullptr);
    if (!workspace || workspace_bytes < need) {
        set_error("lem_time_trace: workspace smaller than lem_time_trace_workspace()");
        return E_WORKSPACE;
    }
    TraceWs L;
    trace_layout(Vs, N, cub_bytes, reinterpret_cast<char *>(workspace), &L);
    hipStream_t st = as_stream(stream);
    if (V > 0) {
        hipLaunchKernelGGL(k_trace_keys, dim3((unsigned)ceil_div(V, TB)), dim3(TB), 0, st, bin, V,
                           N, L.k0, L.v0);
        int rc = check_launch("k_trace_keys");
        if (rc)
            return rc;
        size_t tb = L.cub_bytes;
        if (hipcub::DeviceRadixSort::SortPairs(L.cub, tb, L.k0, L.k1, L.v0, L.v1, (int)V, 0,
                                               key_bits(N - 1), st) != hipSuccess) {
            set_error("lem_time_trace: radix sort failed");
            return E_HIP;
        }
    }
    hipLaunchKernelGGL(k_trace_starts, dim3((unsigned)ceil_div(N + 1, TB)), dim3(TB), 0, st,
                       L.k1, V, N, L.start);
    LEM_REQUIRE(ceil_div(N, TB / 64) <= 2147483647LL, "lem_time_trace: too many samples");
    hipLaunchKernelGGL(k_trace_hist, dim3((unsigned)ceil_div(N, TB / 64)), dim3(TB), 0, st, amp,
                       L.v1, L.start, N, L.H);
    hipLaunchKernelGGL(k_trace_conv, dim3((unsigned)ceil_div(N, TB)), dim3(TB), 0, st, L.H,
                       excitation, E, N, out);
    return check_launch("k_trace_conv");
}
