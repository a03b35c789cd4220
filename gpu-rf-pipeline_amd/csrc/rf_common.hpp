// Shared pieces of the HIP sources: status handling across the C ABI, launch helpers and the
// complex arithmetic every kernel uses. Internal to the library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <exception>
#include <new>
#include <string>

#include "rfhip.h"

namespace rf {

// Thrown inside the library and turned into an rf_status at the C boundary, so no exception
// ever crosses the ABI. Mirrors the reference's throw-on-error style (src/check.hpp:42-66).
struct Error {
    rf_status status;
    std::string message;
};

void set_last_error(const std::string& message);

[[noreturn]] inline void fail(rf_status status, const std::string& message)
{
    throw Error{status, message};
}

[[noreturn]] void fail_hip(hipError_t err, const char* call, const char* file, int line);

#define RF_HIP_CHECK(call)                                                                    \
    do {                                                                                      \
        const hipError_t rf_hip_err_ = (call);                                                \
        if (rf_hip_err_ != hipSuccess) {                                                      \
            ::rf::fail_hip(rf_hip_err_, #call, __FILE__, __LINE__);                           \
        }                                                                                     \
    } while (0)

// Runs `body` and converts a thrown rf::Error (or anything else) into a status code.
template <class F>
rf_status guarded(F&& body)
{
    try {
        body();
        return RF_OK;
    } catch (const Error& e) {
        set_last_error(e.message);
        return e.status;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return RF_ALLOC;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return RF_INVALID_ARG;
    } catch (...) {
        set_last_error("unknown error");
        return RF_INVALID_ARG;
    }
}

inline hipStream_t as_stream(rf_stream s)
{
    return static_cast<hipStream_t>(s);
}

// Checks the launch that was just enqueued.
inline void check_launch(const char* what)
{
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        fail(RF_HIP_ERROR, std::string("HIP error: ") + hipGetErrorString(err) + " (code "
                               + std::to_string(static_cast<int>(err)) + ")\n  in: launch of "
                               + what);
    }
}

// Raises a kernel's dynamic-LDS limit above 64 KiB where the runtime wants that; a runtime
// that does not need it may refuse the call, and the launch itself then reports any problem.
inline void allow_large_lds(const void* kernel, std::size_t bytes)
{
    if (bytes > 64 * 1024
        && hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(bytes)) != hipSuccess) {
        (void)hipGetLastError();
    }
}

// LDS bytes one workgroup may use on the current device.
int max_lds_per_block();
// LDS bytes of one compute unit of the current device.
int lds_per_cu();
// Compute units of the current device.
int cu_count();
// Enqueues the library's streaming device-to-device copy.
void copy_device(const void* src, void* dst, size_t bytes, hipStream_t stream);
// Sets `bytes` bytes at dst to `value` with a kernel (graph-capturable).
void fill_device(void* dst, int value, size_t bytes, hipStream_t stream);

inline long long ceil_div(long long a, long long b)
{
    return (a + b - 1) / b;
}

inline bool is_pow2(long long v)
{
    return v > 0 && (v & (v - 1)) == 0;
}

inline int ilog2(long long v)
{
    int l = 0;
    while ((1LL << l) < v) {
        ++l;
    }
    return l;
}

// Persistent walk over `total` items in XCD-major order: workgroups are dispatched round-robin
// over the 8 XCDs (workgroup b on XCD b % 8), so the items are cut into 8 contiguous ranges and
// range x is walked in order by XCD x's workgroups; neighbouring items, which share cache lines
// (column tiles narrower than a line) or twiddle rows, then run together behind one L2. Needs a
// grid that is a multiple of 8; every item is visited exactly once whatever the XCD count.
template <class Body>
__device__ __forceinline__ void xcd_major_items(int total, Body body)
{
    const int per = (total + 7) >> 3;
    const int lo = (blockIdx.x & 7) * per;
    const int hi = lo + per < total ? lo + per : total;
    const int step = gridDim.x >> 3;
    for (int item = lo + static_cast<int>(blockIdx.x >> 3); item < hi; item += step) {
        body(item);
    }
}

// Device-side complex helpers on float2 (x = re, y = im).
__device__ __forceinline__ float2 cadd(float2 a, float2 b)
{
    return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b)
{
    return make_float2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulconj(float2 a, float2 b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s)
{
    return make_float2(a.x * s, a.y * s);
}
// |a| = sqrt(re^2 + im^2) with one fixed rounding sequence, fma(re, re, im * im): left to
// contraction, the compiler may fuse either product depending on the surrounding code, so the
// magnitude pass and the transforms that store |X| directly would differ in the last bit.
__device__ __forceinline__ float cabs_f(float2 a)
{
    return sqrtf(fmaf(a.x, a.x, a.y * a.y));
}

// Complex loads and stores through the global address space. A pointer read from memory (a
// descriptor struct) is generic to the compiler, which then emits flat instructions; those
// also count against the LDS wait counter, so every LDS wait in a loop would wait for the
// HBM loads in flight too. These keep such accesses global_load / global_store.
typedef __attribute__((address_space(1))) const double global_cdouble;
typedef __attribute__((address_space(1))) double global_double;

__device__ __forceinline__ float2 gload(const float2* p, long long i)
{
    return __builtin_bit_cast(float2, ((global_cdouble*)p)[i]);  // address-space cast
}

__device__ __forceinline__ void gstore(float2* p, long long i, float2 v)
{
    ((global_double*)p)[i] = __builtin_bit_cast(double, v);  // address-space cast
}

}  // namespace rf
