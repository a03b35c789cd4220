// Per-stream workspace blocks (workspace.hpp).

#include "workspace.hpp"

#include <map>
#include <mutex>
#include <thread>
#include <tuple>

namespace rf {
namespace {

// hipStreamPerThread names a different stream in every thread, so its blocks are per thread.
struct StreamKey {
    int device;
    hipStream_t stream;
    std::thread::id thread;

    bool operator<(const StreamKey& o) const
    {
        return std::tie(device, stream, thread) < std::tie(o.device, o.stream, o.thread);
    }
};

std::mutex g_mutex;
std::map<StreamKey, void*> g_blocks;

}  // namespace

void* stream_workspace(hipStream_t stream)
{
    int device = 0;
    RF_HIP_CHECK(hipGetDevice(&device));
    const StreamKey key{device, stream,
                        stream == hipStreamPerThread ? std::this_thread::get_id()
                                                     : std::thread::id()};
    std::lock_guard<std::mutex> lock(g_mutex);
    const auto it = g_blocks.find(key);
    if (it != g_blocks.end()) {
        return it->second;
    }
    hipStreamCaptureMode mode = hipStreamCaptureModeRelaxed;
    RF_HIP_CHECK(hipThreadExchangeStreamCaptureMode(&mode));
    void* block = nullptr;
    const hipError_t err = hipMalloc(&block, kStreamWorkspaceBytes);
    (void)hipThreadExchangeStreamCaptureMode(&mode);
    if (err != hipSuccess) {
        fail_hip(err, "hipMalloc(stream workspace)", __FILE__, __LINE__);
    }
    g_blocks.emplace(key, block);
    return block;
}

void release_stream_workspace(hipStream_t stream)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    for (auto it = g_blocks.begin(); it != g_blocks.end();) {
        if (it->first.stream == stream) {
            (void)hipFree(it->second);
            it = g_blocks.erase(it);
        } else {
            ++it;
        }
    }
}

bool stream_capturing(hipStream_t stream)
{
    hipStreamCaptureStatus status = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &status) != hipSuccess) {
        (void)hipGetLastError();
        return true;
    }
    return status != hipStreamCaptureStatusNone;
}

}  // namespace rf
