// Internal helpers shared by the oppx HIP translation units (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "oppx.h"

namespace oppx {

void set_error(const std::string& msg);

// Process-wide launch shape (oppx_set_tuning), defined in oppx_runtime.hip.
extern oppx_tuning g_tuning;

// Device address (*out, nullable) of the static scratch of the allocation-free entry point on `device`,
// which must be current; looked up once per device. Defined in oppx_compact.hip, called by oppx_init.
int noalloc_scratch(int device, void** out);

// Blocks of `kernel` (with `block` threads) resident on the whole device; the occupancy query is
// cached per (kernel, block).
unsigned resident_blocks(const void* kernel, int block);

// Returns from the enclosing C-ABI function with OPPX_ERR_HIP when a HIP call fails.
#define OPPX_HIP_TRY(call)                                                                       \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            ::oppx::set_error(std::string(#call) + ": " + hipGetErrorString(e_));                 \
            return OPPX_ERR_HIP;                                                                \
        }                                                                                       \
    } while (0)

#define OPPX_ARG_CHECK(cond, msg)                                                                \
    do {                                                                                        \
        if (!(cond)) {                                                                          \
            ::oppx::set_error(msg);                                                             \
            return OPPX_ERR_ARG;                                                                \
        }                                                                                       \
    } while (0)

// Launch-error check right after a <<<>>> launch.
#define OPPX_LAUNCH_CHECK(name)                                                                  \
    do {                                                                                        \
        hipError_t e_ = hipGetLastError();                                                      \
        if (e_ != hipSuccess) {                                                                 \
            ::oppx::set_error(std::string("launch ") + name + ": " + hipGetErrorString(e_));     \
            return OPPX_ERR_HIP;                                                                \
        }                                                                                       \
    } while (0)

constexpr int kWave = 64;
constexpr int kMaxCell = OPPX_MAX_CELL_INDEX;
constexpr int kXcds = 8;  // MI355X: 8 XCDs, workgroups dispatched round-robin across them

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Workgroups are dispatched to XCDs round-robin (block b runs on XCD b % 8). Remap so that
// each XCD processes one contiguous 1/8 of the logical blocks: neighbouring particles then
// share one XCD's L2 for their mesh gathers.
__device__ __forceinline__ uint32_t xcd_block(uint32_t pid, uint32_t nblocks) {
    const uint32_t q = nblocks / kXcds, r = nblocks % kXcds;
    const uint32_t x = pid % kXcds, k = pid / kXcds;
    return x * q + (x < r ? x : r) + k;
}

// Blocks of per_block particles that hold live particles (n of them), at most nblocks. A grid sized by
// the host's upper bound on the live count lets workgroups 0 .. live_blocks - 1, which the hardware
// dispatches first, take those blocks through xcd_block(pid, live_blocks): every XCD then gets 1/8 of the
// live particles however far the count has fallen below the bound, and the remaining workgroups leave.
__device__ __forceinline__ uint32_t live_blocks(int64_t n, int64_t per_block, uint32_t nblocks) {
    const int64_t l = (n + per_block - 1) / per_block;
    return l < (int64_t)nblocks ? (uint32_t)l : nblocks;
}

__device__ __forceinline__ int wave_sum_i32(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ long long wave_sum_i64(long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int w = __shfl_xor(v, o);
        v = v > w ? v : w;
    }
    return v;
}

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// Grid of a kernel whose blocks stride over `units` units of work (one block each): at most `mult`
// times the blocks of `kernel` that fit the device at once. Units past a live count that the kernel
// reads on the device then cost no dispatched blocks; each dead block would otherwise still start its
// waves and wait for that read before leaving.
inline unsigned stride_grid(const void* kernel, int block, int64_t units, unsigned mult) {
    const int64_t cap = (int64_t)resident_blocks(kernel, block) * mult;
    const int64_t g = units < cap ? units : cap;
    return (unsigned)(g > 0 ? g : 1);
}

}  // namespace oppx
