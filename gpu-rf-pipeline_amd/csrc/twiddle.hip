// Twiddle tables (float64 on the host, stored as float32 on the device).

#include "twiddle.hpp"
#include "fft16384.hpp"

#include <cmath>
#include <mutex>
#include <vector>

namespace rf {

void fill_twiddles(float2* host, long long count, long long n, long long step)
{
    // angle = -2 pi (m * step mod n) / n, reduced exactly in integers before the float64
    // sin/cos so that large indices keep full accuracy.
    const double two_pi = 6.283185307179586476925286766559;
    for (long long m = 0; m < count; ++m) {
        const long long e = (m * step) % n;
        const double a = -two_pi * static_cast<double>(e) / static_cast<double>(n);
        host[m] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
    }
}

namespace {
constexpr int kMaxDevices = 64;
std::once_flag g_once[kMaxDevices];
float2* g_table[kMaxDevices] = {};
}  // namespace

const float2* twiddle_table_4096()
{
    int device = 0;
    RF_HIP_CHECK(hipGetDevice(&device));
    if (device < 0 || device >= kMaxDevices) {
        fail(RF_LIMIT, "twiddle table: device index out of range");
    }
    std::call_once(g_once[device], [device] {
        std::vector<float2> host(kTwiddleSize);
        fill_twiddles(host.data(), kTwiddleSize, kTwiddleSize, 1);
        float2* dev = nullptr;
        RF_HIP_CHECK(hipMalloc(&dev, kTwiddleSize * sizeof(float2)));
        RF_HIP_CHECK(hipMemcpy(dev, host.data(), kTwiddleSize * sizeof(float2),
                               hipMemcpyHostToDevice));
        g_table[device] = dev;
    });
    if (g_table[device] == nullptr) {
        fail(RF_HIP_ERROR, "twiddle table: initialisation failed on an earlier call");
    }
    return g_table[device];
}

namespace {
std::once_flag g_stage_once[kMaxDevices];
float2* g_stage[kMaxDevices] = {};
}  // namespace

const float2* fft4096_stage_table()
{
    int device = 0;
    RF_HIP_CHECK(hipGetDevice(&device));
    if (device < 0 || device >= kMaxDevices) {
        fail(RF_LIMIT, "twiddle table: device index out of range");
    }
    std::call_once(g_stage_once[device], [device] {
        std::vector<float2> host;
        for (int span : {16, 256}) {
            for (int k = 0; k < span; ++k) {
                float2 w[16];
                // exp(-2 pi i r k / (16 span)): entries r*k of a transform of 16*span points
                fill_twiddles(w, 16, 16LL * span, k);
                host.insert(host.end(), w, w + 16);
            }
        }
        float2* dev = nullptr;
        RF_HIP_CHECK(hipMalloc(&dev, host.size() * sizeof(float2)));
        RF_HIP_CHECK(hipMemcpy(dev, host.data(), host.size() * sizeof(float2),
                               hipMemcpyHostToDevice));
        g_stage[device] = dev;
    });
    if (g_stage[device] == nullptr) {
        fail(RF_HIP_ERROR, "twiddle table: initialisation failed on an earlier call");
    }
    return g_stage[device];
}

namespace {
std::once_flag g_16384_once[kMaxDevices];
float2* g_16384[kMaxDevices] = {};
}  // namespace

const float2* fft16384_table()
{
    int device = 0;
    RF_HIP_CHECK(hipGetDevice(&device));
    if (device < 0 || device >= kMaxDevices) {
        fail(RF_LIMIT, "twiddle table: device index out of range");
    }
    std::call_once(g_16384_once[device], [device] {
        std::vector<float2> host(kTw16384Size, make_float2(0.0f, 0.0f));
        for (int k = 0; k < 4; ++k) {
            fill_twiddles(host.data() + kTw4 + 16 * k, 16, 64, k);      // exp(-2 pi i r k / 64)
        }
        for (int k = 0; k < 64; ++k) {
            fill_twiddles(host.data() + kTw64 + 17 * k, 16, 1024, k);   // exp(-2 pi i r k / 1024)
        }
        fill_twiddles(host.data() + kTwHi, 128, 128, 1);
        fill_twiddles(host.data() + kTwLo, 128, 16384, 1);
        fill_twiddles(host.data() + kFoldLane, 1024, 32768, 1);
        fill_twiddles(host.data() + kFoldRow, 32, 64, 1);
        float2* dev = nullptr;
        RF_HIP_CHECK(hipMalloc(&dev, host.size() * sizeof(float2)));
        RF_HIP_CHECK(hipMemcpy(dev, host.data(), host.size() * sizeof(float2),
                               hipMemcpyHostToDevice));
        g_16384[device] = dev;
    });
    if (g_16384[device] == nullptr) {
        fail(RF_HIP_ERROR, "twiddle table: initialisation failed on an earlier call");
    }
    return g_16384[device];
}

}  // namespace rf
