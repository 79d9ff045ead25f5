// Device runtime of the oppx C-ABI: initialisation, errors, memory, AoS<->SoA transposes.
// Replaces opp_lib/src/hip/opp_hip.cpp:221-271 (device init) and the host-side layout
// conversions of opp_lib/src/hip/opp_hip_utils.cpp:35-153, which here run on the device.

#include <cstring>

#include "oppx_internal.h"

namespace oppx {

static thread_local std::string g_last_error;

oppx_tuning g_tuning = {256, 16, 0, 0, 16, 64, 256, 1};

unsigned resident_blocks(const void* kernel, int block) {
    static int cus = 0;
    static const void* kern_cache[128] = {nullptr};
    static int block_cache[128] = {0}, per_cu_cache[128] = {0};
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    int per_cu = 0;
    for (int i = 0; i < 128; i++) {
        if (kern_cache[i] == kernel && block_cache[i] == block) { per_cu = per_cu_cache[i]; break; }
        if (kern_cache[i] == nullptr) {
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu <= 0)
                per_cu = 1;
            kern_cache[i] = kernel;
            block_cache[i] = block;
            per_cu_cache[i] = per_cu;
            break;
        }
    }
    return (unsigned)cus * (unsigned)(per_cu > 0 ? per_cu : 1);
}

void set_error(const std::string& msg) { g_last_error = msg; }

namespace {

// One thread per (element, component); reads/writes coalesced on the SoA side and
// contiguous on the AoS side.
template <typename T>
__global__ void aos_to_soa_kernel(const T* __restrict__ aos, T* __restrict__ soa, int64_t n, int dim,
                                  int64_t stride) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * dim) return;
    const int64_t i = t / dim, d = t % dim;
    soa[d * stride + i] = aos[t];
}

template <typename T>
__global__ void soa_to_aos_kernel(const T* __restrict__ soa, T* __restrict__ aos, int64_t n, int dim,
                                  int64_t stride) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * dim) return;
    const int64_t i = t / dim, d = t % dim;
    aos[t] = soa[d * stride + i];
}

struct Pattern {
    unsigned char b[256];
};

// element i, component d = pattern[d]; AoS when stride == 0, else SoA with the given stride
template <typename T>
__global__ void fill_pattern_kernel(T* __restrict__ dst, int64_t n, int dim, int64_t stride, Pattern pat) {
    // one record per thread and step, its components in an inner loop: no 64-bit division per element
    const T* pv = reinterpret_cast<const T*>(pat.b);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (stride == 0) {
            for (int d = 0; d < dim; d++) dst[i * dim + d] = pv[d];  // AoS: record i is dim consecutive values
        } else {
            for (int d = 0; d < dim; d++) dst[d * stride + i] = pv[d];  // SoA: component d at d * stride + i
        }
    }
}

// entries [first, first + count) of every `rec`-int record: v -> v + offset unless v < 0 or v == MAX
__global__ void offset_indices_kernel(int32_t* __restrict__ base, int64_t n, int rec, int first, int count,
                                      int32_t offset) {
    // one record per thread and step (no 64-bit division per entry)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int32_t* p = base + i * rec + first;
        for (int k = 0; k < count; k++) {
            const int32_t v = p[k];
            if (v >= 0 && v != kMaxCell) p[k] = v + offset;
        }
    }
}

}  // namespace
}  // namespace oppx

using namespace oppx;

extern "C" {

const char* oppx_last_error(void) { return g_last_error.c_str(); }

int oppx_init(int device) {
    int count = 0;
    OPPX_HIP_TRY(hipGetDeviceCount(&count));
    OPPX_ARG_CHECK(count > 0, "oppx_init: no HIP device visible");
    OPPX_ARG_CHECK(device >= 0 && device < count, "oppx_init: device index out of range");
    OPPX_HIP_TRY(hipSetDevice(device));
    return noalloc_scratch(device, nullptr);
}

int oppx_device_malloc(void** ptr, size_t bytes) {
    OPPX_ARG_CHECK(ptr != nullptr, "oppx_device_malloc: null out pointer");
    *ptr = nullptr;
    if (bytes == 0) return OPPX_OK;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
        *ptr = nullptr;
        return OPPX_ERR_NOMEM;
    }
    return OPPX_OK;
}

int oppx_device_free(void* ptr) {
    if (ptr) OPPX_HIP_TRY(hipFree(ptr));
    return OPPX_OK;
}

int oppx_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return OPPX_OK;
    OPPX_ARG_CHECK(dst && src, "oppx_memcpy_htod: null pointer");
    OPPX_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    OPPX_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return OPPX_OK;
}

int oppx_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return OPPX_OK;
    OPPX_ARG_CHECK(dst && src, "oppx_memcpy_dtoh: null pointer");
    OPPX_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    OPPX_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return OPPX_OK;
}

int oppx_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return OPPX_OK;
    OPPX_ARG_CHECK(dst && src, "oppx_memcpy_dtod: null pointer");
    OPPX_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
    return OPPX_OK;
}

int oppx_memset_zero(void* dst, size_t bytes, void* stream) {
    if (bytes == 0) return OPPX_OK;
    OPPX_ARG_CHECK(dst, "oppx_memset_zero: null pointer");
    OPPX_HIP_TRY(hipMemsetAsync(dst, 0, bytes, as_stream(stream)));
    return OPPX_OK;
}

int oppx_stream_sync(void* stream) {
    OPPX_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return OPPX_OK;
}

int oppx_aos_to_soa(const void* aos, void* soa, int64_t n, int dim, int64_t stride, int elem_bytes, void* stream) {
    OPPX_ARG_CHECK(n >= 0 && dim > 0 && stride >= n, "oppx_aos_to_soa: bad sizes");
    OPPX_ARG_CHECK(elem_bytes == 4 || elem_bytes == 8, "oppx_aos_to_soa: elem_bytes must be 4 or 8");
    if (n == 0) return OPPX_OK;
    OPPX_ARG_CHECK(aos && soa, "oppx_aos_to_soa: null pointer");
    const unsigned g = grid_for(n * dim, 256);
    if (elem_bytes == 8)
        aos_to_soa_kernel<double><<<g, 256, 0, as_stream(stream)>>>((const double*)aos, (double*)soa, n, dim, stride);
    else
        aos_to_soa_kernel<int32_t><<<g, 256, 0, as_stream(stream)>>>((const int32_t*)aos, (int32_t*)soa, n, dim, stride);
    OPPX_LAUNCH_CHECK("aos_to_soa");
    return OPPX_OK;
}

int oppx_soa_to_aos(const void* soa, void* aos, int64_t n, int dim, int64_t stride, int elem_bytes, void* stream) {
    OPPX_ARG_CHECK(n >= 0 && dim > 0 && stride >= n, "oppx_soa_to_aos: bad sizes");
    OPPX_ARG_CHECK(elem_bytes == 4 || elem_bytes == 8, "oppx_soa_to_aos: elem_bytes must be 4 or 8");
    if (n == 0) return OPPX_OK;
    OPPX_ARG_CHECK(aos && soa, "oppx_soa_to_aos: null pointer");
    const unsigned g = grid_for(n * dim, 256);
    if (elem_bytes == 8)
        soa_to_aos_kernel<double><<<g, 256, 0, as_stream(stream)>>>((const double*)soa, (double*)aos, n, dim, stride);
    else
        soa_to_aos_kernel<int32_t><<<g, 256, 0, as_stream(stream)>>>((const int32_t*)soa, (int32_t*)aos, n, dim, stride);
    OPPX_LAUNCH_CHECK("soa_to_aos");
    return OPPX_OK;
}

int oppx_fill_pattern(void* dst, int64_t n, int dim, int elem_bytes, const void* pattern, int64_t stride, void* stream) {
    OPPX_ARG_CHECK(n >= 0 && dim > 0 && (elem_bytes == 4 || elem_bytes == 8), "oppx_fill_pattern: bad sizes");
    OPPX_ARG_CHECK((int64_t)dim * elem_bytes <= 256, "oppx_fill_pattern: pattern above 256 bytes");
    OPPX_ARG_CHECK(stride == 0 || stride >= n, "oppx_fill_pattern: stride below n");
    if (n == 0) return OPPX_OK;
    OPPX_ARG_CHECK(dst && pattern, "oppx_fill_pattern: null pointer");
    Pattern p{};
    memcpy(p.b, pattern, (size_t)dim * elem_bytes);
    int64_t g = grid_for(n, 256);
    if (g > 8192) g = 8192;
    if (elem_bytes == 8)
        fill_pattern_kernel<double><<<(unsigned)g, 256, 0, as_stream(stream)>>>((double*)dst, n, dim, stride, p);
    else
        fill_pattern_kernel<int32_t><<<(unsigned)g, 256, 0, as_stream(stream)>>>((int32_t*)dst, n, dim, stride, p);
    OPPX_LAUNCH_CHECK("fill_pattern");
    return OPPX_OK;
}

int oppx_offset_indices(int32_t* base, int64_t n_records, int rec_ints, int first, int count, int32_t offset,
                        void* stream) {
    OPPX_ARG_CHECK(n_records >= 0 && rec_ints > 0 && first >= 0 && count > 0 && first + count <= rec_ints,
                   "oppx_offset_indices: bad record layout");
    if (n_records == 0 || offset == 0) return OPPX_OK;
    OPPX_ARG_CHECK(base != nullptr, "oppx_offset_indices: null pointer");
    int64_t g = grid_for(n_records, 256);
    if (g > 8192) g = 8192;
    offset_indices_kernel<<<(unsigned)g, 256, 0, as_stream(stream)>>>(base, n_records, rec_ints, first, count, offset);
    OPPX_LAUNCH_CHECK("offset_indices");
    return OPPX_OK;
}

}  // extern "C"
