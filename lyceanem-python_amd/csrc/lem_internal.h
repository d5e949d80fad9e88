// lem_internal.h -- host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace lem {

// Thread-local last-error message (lem_last_error()).
void set_error(const std::string &msg);
void clear_error();

// Status codes (mirrored in include/lem.h).
constexpr int OK = 0;
constexpr int E_INVALID = -1;
constexpr int E_HIP = -2;
constexpr int E_WORKSPACE = -3;

// Checks the launch that was just issued; records the HIP error text.
int check_launch(const char *what);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Source-sum machinery of lem_fields.hip, shared with the grid-fused kernel.
int64_t los_chunk_count(int64_t S, int64_t K);
int launch_los_finish(const float *ws_sum, const float *ws_cmp, int64_t NC, int64_t K,
                      const float *snk_n, const float *snk_w, int tangent, float *out,
                      hipStream_t st);

// Spatial (3-D Morton) order of n points, computed on the device (lem_grid.hip):
// point_order_workspace(n) bytes of workspace (256-byte aligned); *order is set
// to the int32 order[n] inside it.  Returns -1 from the size query on failure.
int64_t point_order_workspace(int64_t n);
int point_order(const float *pts, int64_t n, void *ws, const int32_t **order, hipStream_t st);

}  // namespace lem

#define LEM_REQUIRE(cond, msg)            \
    do {                                  \
        if (!(cond)) {                    \
            ::lem::set_error(msg);        \
            return ::lem::E_INVALID;      \
        }                                 \
    } while (0)
