#pragma once

// Owning device allocation through the C ABI (reference: src/device_buffer.hpp:11-76).

#include "check.hpp"

#include <cstddef>
#include <cstdio>

template <class T>
class DeviceBuffer {
public:
    explicit DeviceBuffer(std::size_t count) : count_(count)
    {
        if (count_ > 0) {
            void* raw = nullptr;
            RF_CHECK(rf_malloc(&raw, bytes()));
            ptr_ = static_cast<T*>(raw);
        }
    }
    ~DeviceBuffer()
    {
        if (ptr_ != nullptr && rf_free(ptr_) != RF_OK) {
            std::fprintf(stderr, "warning: rf_free failed: %s\n", rf_last_error());
        }
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;

    T* get() const { return ptr_; }
    std::size_t size() const { return count_; }
    std::size_t bytes() const { return count_ * sizeof(T); }

    // Host-to-device copy into elements [offset, offset + count); returns when it is done.
    void upload(const T* src, std::size_t count, std::size_t offset = 0)
    {
        if (count > 0) {
            RF_CHECK(rf_memcpy_h2d(ptr_ + offset, src, count * sizeof(T), nullptr));
            RF_CHECK(rf_stream_synchronize(nullptr));
        }
    }
    void upload_at(std::size_t offset, const T* src, std::size_t count)
    {
        upload(src, count, offset);
    }
    // Device-to-host copy of the first `count` elements, after the work queued before it.
    void download(T* dst, std::size_t count) const
    {
        if (count > 0) {
            RF_CHECK(rf_memcpy_d2h(dst, ptr_, count * sizeof(T), nullptr));
            RF_CHECK(rf_stream_synchronize(nullptr));
        }
    }
    void zero()
    {
        if (count_ > 0) {
            RF_CHECK(rf_memset(ptr_, 0, bytes(), nullptr));
        }
    }

private:
    T* ptr_ = nullptr;
    std::size_t count_ = 0;
};
