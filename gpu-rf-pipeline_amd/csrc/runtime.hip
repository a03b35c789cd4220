// Status codes, the thread-local error text, and the runtime utilities of the C ABI (device
// memory, streams, events). These have no arithmetic; the reference calls the CUDA runtime for
// the same jobs directly (src/device_buffer.hpp:11-76, src/bench.cpp:10-40).

#include "rf_common.hpp"
#include "workspace.hpp"

#include <cstring>
#include <mutex>
#include <string>

namespace rf {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& message)
{
    g_last_error = message;
}

void fail_hip(hipError_t err, const char* call, const char* file, int line)
{
    fail(err == hipErrorOutOfMemory || err == hipErrorMemoryAllocation ? RF_ALLOC : RF_HIP_ERROR,
         std::string("HIP error: ") + hipGetErrorString(err) + " (code "
             + std::to_string(static_cast<int>(err)) + ")\n  in: " + call + "\n  at: " + file
             + ":" + std::to_string(line));
}

namespace {

// The device limits the launch paths consult, queried once per device: a FIR or channelizer
// call over a small capture is a few microseconds of GPU work, and three attribute queries per
// call would add host time of the same order.
struct DeviceLimits {
    int lds_per_block = 0;
    int lds_per_cu = 0;
    int cus = 0;
};

constexpr int kMaxLimitDevices = 64;
std::once_flag g_limits_once[kMaxLimitDevices];
DeviceLimits g_limits[kMaxLimitDevices];

const DeviceLimits& device_limits()
{
    int device = 0;
    RF_HIP_CHECK(hipGetDevice(&device));
    if (device < 0 || device >= kMaxLimitDevices) {
        fail(RF_LIMIT, "device index " + std::to_string(device) + " out of range");
    }
    std::call_once(g_limits_once[device], [device] {
        DeviceLimits& l = g_limits[device];
        RF_HIP_CHECK(hipDeviceGetAttribute(&l.lds_per_block,
                                           hipDeviceAttributeMaxSharedMemoryPerBlock, device));
        RF_HIP_CHECK(hipDeviceGetAttribute(&l.lds_per_cu,
                                           hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                           device));
        RF_HIP_CHECK(hipDeviceGetAttribute(&l.cus, hipDeviceAttributeMultiprocessorCount, device));
    });
    return g_limits[device];
}

}  // namespace

int max_lds_per_block()
{
    return device_limits().lds_per_block;
}

int lds_per_cu()
{
    return device_limits().lds_per_cu;
}

int cu_count()
{
    return device_limits().cus;
}

namespace {

// Device-to-device copy as a streaming kernel: 16-byte loads and stores when both pointers
// allow it, a grid of a few workgroups per CU striding over the buffer. It is also the
// reference point for the HBM-bandwidth and traffic-counter calibration.
__global__ __launch_bounds__(256) void copy_kernel(const unsigned char* __restrict__ src,
                                                   unsigned char* __restrict__ dst, size_t bytes)
{
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    const size_t first = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool vec = ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15) == 0;
    size_t done = 0;
    if (vec) {
        // Four 16-byte loads in flight per lane before their stores.
        const size_t n16 = bytes / 16;
        const float4* s4 = reinterpret_cast<const float4*>(src);
        float4* d4 = reinterpret_cast<float4*>(dst);
        size_t i = first;
        for (; i + 3 * stride < n16; i += 4 * stride) {
            const float4 a = s4[i];
            const float4 b = s4[i + stride];
            const float4 c = s4[i + 2 * stride];
            const float4 d = s4[i + 3 * stride];
            d4[i] = a;
            d4[i + stride] = b;
            d4[i + 2 * stride] = c;
            d4[i + 3 * stride] = d;
        }
        for (; i < n16; i += stride) {
            d4[i] = s4[i];
        }
        done = n16 * 16;
    }
    for (size_t i = done + first; i < bytes; i += stride) {
        dst[i] = src[i];
    }
}

// Byte fill for rf_memset: a kernel rather than hipMemsetAsync, so a stream capture records
// an ordinary kernel node (memset nodes have misbehaved on graph replays on this platform).
// 16-byte stores for the aligned body, bytes for the unaligned head and tail.
__global__ __launch_bounds__(256) void fill_kernel(unsigned char* __restrict__ dst,
                                                   unsigned char value, size_t head, size_t n16,
                                                   size_t bytes)
{
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    const size_t first = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const unsigned v4 = 0x01010101u * value;
    const uint4 word = make_uint4(v4, v4, v4, v4);
    uint4* body = reinterpret_cast<uint4*>(dst + head);
    for (size_t i = first; i < n16; i += stride) {
        body[i] = word;
    }
    const size_t tail = head + n16 * 16;
    for (size_t i = first; i < head; i += stride) {
        dst[i] = value;
    }
    for (size_t i = tail + first; i < bytes; i += stride) {
        dst[i] = value;
    }
}

}  // namespace

void fill_device(void* dst, int value, size_t bytes, hipStream_t stream)
{
    if (bytes == 0) {
        return;
    }
    unsigned char* d = static_cast<unsigned char*>(dst);
    size_t head = (16 - (reinterpret_cast<uintptr_t>(d) & 15)) & 15;
    head = head < bytes ? head : bytes;
    const size_t n16 = (bytes - head) / 16;
    const long long cap = 8LL * cu_count();
    long long grid = ceil_div(static_cast<long long>(n16 + 1), 256);
    grid = grid < cap ? grid : cap;
    fill_kernel<<<static_cast<unsigned>(grid), 256, 0, stream>>>(
        d, static_cast<unsigned char>(value), head, n16, bytes);
    check_launch("fill_kernel");
}

void copy_device(const void* src, void* dst, size_t bytes, hipStream_t stream)
{
    if (bytes == 0) {
        return;
    }
    const long long cap = 8LL * cu_count();
    long long grid = ceil_div(static_cast<long long>(bytes / 16 + 1), 256);
    grid = grid < cap ? grid : cap;
    copy_kernel<<<static_cast<unsigned>(grid), 256, 0, stream>>>(
        static_cast<const unsigned char*>(src), static_cast<unsigned char*>(dst), bytes);
    check_launch("copy_kernel");
}

}  // namespace rf

using rf::guarded;

extern "C" {

const char* rf_status_string(rf_status status)
{
    switch (status) {
    case RF_OK:          return "RF_OK";
    case RF_INVALID_ARG: return "RF_INVALID_ARG";
    case RF_LIMIT:       return "RF_LIMIT";
    case RF_HIP_ERROR:   return "RF_HIP_ERROR";
    case RF_ALLOC:       return "RF_ALLOC";
    }
    return "RF_UNKNOWN_STATUS";
}

const char* rf_last_error(void)
{
    return rf::g_last_error.c_str();
}

int rf_abi_version(void)
{
    return 1;
}

rf_status rf_device_count(int* count)
{
    return guarded([&] {
        if (count == nullptr) rf::fail(RF_INVALID_ARG, "rf_device_count: count is NULL");
        *count = 0;
        RF_HIP_CHECK(hipGetDeviceCount(count));
    });
}

rf_status rf_device_name(char* buffer, int length)
{
    return guarded([&] {
        if (buffer == nullptr || length < 1) {
            rf::fail(RF_INVALID_ARG, "rf_device_name: need a buffer");
        }
        int device = 0;
        RF_HIP_CHECK(hipGetDevice(&device));
        hipDeviceProp_t prop{};
        RF_HIP_CHECK(hipGetDeviceProperties(&prop, device));
        std::string name = prop.name;
        if (name.empty()) name = prop.gcnArchName;
        std::strncpy(buffer, name.c_str(), static_cast<size_t>(length) - 1);
        buffer[length - 1] = '\0';
    });
}

rf_status rf_device_memory(size_t* bytes)
{
    return guarded([&] {
        int device = 0;
        RF_HIP_CHECK(hipGetDevice(&device));
        hipDeviceProp_t prop{};
        RF_HIP_CHECK(hipGetDeviceProperties(&prop, device));
        *bytes = prop.totalGlobalMem;
    });
}

rf_status rf_malloc(void** ptr, size_t bytes)
{
    return guarded([&] {
        if (ptr == nullptr) rf::fail(RF_INVALID_ARG, "rf_malloc: ptr is NULL");
        *ptr = nullptr;
        if (bytes == 0) return;
        RF_HIP_CHECK(hipMalloc(ptr, bytes));
    });
}

rf_status rf_free(void* ptr)
{
    return guarded([&] { RF_HIP_CHECK(hipFree(ptr)); });
}

rf_status rf_host_alloc(void** ptr, size_t bytes)
{
    return guarded([&] {
        if (ptr == nullptr) rf::fail(RF_INVALID_ARG, "rf_host_alloc: ptr is NULL");
        *ptr = nullptr;
        if (bytes == 0) return;
        RF_HIP_CHECK(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    });
}

rf_status rf_host_alloc_mapped(void** host, void** device, size_t bytes)
{
    return guarded([&] {
        if (host == nullptr || device == nullptr) {
            rf::fail(RF_INVALID_ARG, "rf_host_alloc_mapped: NULL output");
        }
        *host = nullptr;
        *device = nullptr;
        if (bytes == 0) return;
        RF_HIP_CHECK(hipHostMalloc(host, bytes, hipHostMallocMapped));
        RF_HIP_CHECK(hipHostGetDevicePointer(device, *host, 0));
    });
}

rf_status rf_host_free(void* ptr)
{
    return guarded([&] { RF_HIP_CHECK(hipHostFree(ptr)); });
}

rf_status rf_memcpy_h2d(void* dst, const void* src, size_t bytes, rf_stream stream)
{
    return guarded([&] {
        if (bytes == 0) return;
        RF_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, rf::as_stream(stream)));
    });
}

rf_status rf_memcpy_d2h(void* dst, const void* src, size_t bytes, rf_stream stream)
{
    return guarded([&] {
        if (bytes == 0) return;
        RF_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, rf::as_stream(stream)));
    });
}

rf_status rf_memcpy_d2d(void* dst, const void* src, size_t bytes, rf_stream stream)
{
    return guarded([&] {
        if (bytes == 0) return;
        if (dst == nullptr || src == nullptr) rf::fail(RF_INVALID_ARG, "rf_memcpy_d2d: NULL");
        rf::copy_device(src, dst, bytes, rf::as_stream(stream));
    });
}

rf_status rf_memset(void* dst, int value, size_t bytes, rf_stream stream)
{
    return guarded([&] {
        if (bytes == 0) return;
        if (dst == nullptr) rf::fail(RF_INVALID_ARG, "rf_memset: NULL");
        rf::fill_device(dst, value, bytes, rf::as_stream(stream));
    });
}

rf_status rf_stream_create(rf_stream* stream)
{
    return guarded([&] {
        hipStream_t s = nullptr;
        RF_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        try {
            // the stream's workspace block now, so no later call on it allocates
            (void)rf::stream_workspace(s);
        } catch (...) {
            (void)hipStreamDestroy(s);
            throw;
        }
        *stream = s;
    });
}

rf_status rf_stream_destroy(rf_stream stream)
{
    return guarded([&] {
        hipStream_t s = rf::as_stream(stream);
        // the stream's workspace block (if any) is freed once the stream's work has finished
        RF_HIP_CHECK(hipStreamSynchronize(s));
        rf::release_stream_workspace(s);
        RF_HIP_CHECK(hipStreamDestroy(s));
    });
}

rf_status rf_stream_synchronize(rf_stream stream)
{
    return guarded([&] { RF_HIP_CHECK(hipStreamSynchronize(rf::as_stream(stream))); });
}

rf_status rf_device_synchronize(void)
{
    return guarded([&] { RF_HIP_CHECK(hipDeviceSynchronize()); });
}

rf_status rf_event_create(void** event)
{
    return guarded([&] {
        hipEvent_t e = nullptr;
        RF_HIP_CHECK(hipEventCreate(&e));
        *event = e;
    });
}

rf_status rf_event_destroy(void* event)
{
    return guarded([&] { RF_HIP_CHECK(hipEventDestroy(static_cast<hipEvent_t>(event))); });
}

rf_status rf_event_record(void* event, rf_stream stream)
{
    return guarded([&] {
        RF_HIP_CHECK(hipEventRecord(static_cast<hipEvent_t>(event), rf::as_stream(stream)));
    });
}

rf_status rf_stream_wait_event(rf_stream stream, void* event)
{
    return guarded([&] {
        if (event == nullptr) rf::fail(RF_INVALID_ARG, "stream wait: NULL event");
        RF_HIP_CHECK(hipStreamWaitEvent(rf::as_stream(stream), static_cast<hipEvent_t>(event), 0));
    });
}

rf_status rf_event_synchronize(void* event)
{
    return guarded([&] { RF_HIP_CHECK(hipEventSynchronize(static_cast<hipEvent_t>(event))); });
}

rf_status rf_event_elapsed_ms(void* start, void* stop, float* ms)
{
    return guarded([&] {
        RF_HIP_CHECK(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start),
                                         static_cast<hipEvent_t>(stop)));
    });
}

rf_status rf_graph_begin_capture(rf_stream stream)
{
    return guarded([&] {
        if (stream == nullptr) rf::fail(RF_INVALID_ARG, "graph capture: needs a created stream");
        RF_HIP_CHECK(hipStreamBeginCapture(rf::as_stream(stream), hipStreamCaptureModeThreadLocal));
    });
}

rf_status rf_graph_end_capture(rf_stream stream, void** graph)
{
    return guarded([&] {
        if (stream == nullptr || graph == nullptr) rf::fail(RF_INVALID_ARG, "graph capture: NULL");
        *graph = nullptr;
        hipGraph_t g = nullptr;
        RF_HIP_CHECK(hipStreamEndCapture(rf::as_stream(stream), &g));
        hipGraphExec_t exec = nullptr;
        const hipError_t err = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (err != hipSuccess) {
            rf::fail_hip(err, "hipGraphInstantiate", __FILE__, __LINE__);
        }
        *graph = exec;
    });
}

rf_status rf_graph_launch(void* graph, rf_stream stream)
{
    return guarded([&] {
        if (graph == nullptr) rf::fail(RF_INVALID_ARG, "graph launch: NULL graph");
        RF_HIP_CHECK(hipGraphLaunch(static_cast<hipGraphExec_t>(graph), rf::as_stream(stream)));
    });
}

rf_status rf_graph_destroy(void* graph)
{
    return guarded([&] {
        if (graph != nullptr) {
            RF_HIP_CHECK(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph)));
        }
    });
}

}  // extern "C"
