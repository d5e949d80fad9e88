// lem_host_shim.h -- host stand-ins for the few HIP pieces that lem_device.h and
// lem_grid_walk.h use, so the grid walk and the grid build's per-item
// arithmetic compile as plain C++ for the CPU checks in tests/ (test
// infrastructure; the product never includes this file).
#pragma once

#define LEM_HOST_BUILD 1
#define LEM_DEV inline

#include <algorithm>
#include <cmath>
#include <cstdint>

struct float4 {
    float x, y, z, w;
};
struct uint2 {
    unsigned int x, y;
};
inline float4 make_float4(float x, float y, float z, float w) { return {x, y, z, w}; }
inline uint2 make_uint2(unsigned int x, unsigned int y) { return {x, y}; }

namespace lem {
using std::max;
using std::min;
// The device code refines or tolerates these hardware estimates; on the host
// the exact operation stands in (recip_f32_to_f64's two Newton steps leave a
// correctly rounded reciprocal unchanged, so F is the same double).
inline double rcp_estimate(double a) { return 1.0 / a; }
inline float rcpf_estimate(float a) { return 1.0f / a; }
inline float sqrtf_estimate(float a) { return std::sqrt(a); }
}  // namespace lem
