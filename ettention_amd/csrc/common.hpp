// Error handling and device-memory helpers shared by the host code of libetn.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

#include "etn/etn.h"

namespace etn {

// Exceptions used inside the library; they never cross the C ABI (capi.cpp
// maps them to etn_status).  Mirrors the reference's Exception / GPUException
// split (ettention/libReconstruction/src/framework/error/Exception.h,
// src/gpu/opencl/OpenCLHeader.h:10-16).
struct Error : std::runtime_error {
    etn_status status;
    Error(etn_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

// Stores the message etn_last_error() returns on this thread; returns s.
etn_status record_error(etn_status s, const std::string& message);

[[noreturn]] inline void fail(const std::string& m) { throw Error(ETN_ERROR_INVALID_ARGUMENT, m); }
[[noreturn]] inline void unsupported(const std::string& m) { throw Error(ETN_ERROR_UNSUPPORTED, m); }

inline void hip_check(hipError_t e, const char* what)
{
    if (e == hipSuccess) return;
    etn_status s = (e == hipErrorOutOfMemory) ? ETN_ERROR_OUT_OF_MEMORY : ETN_ERROR_DEVICE;
    throw Error(s, std::string(what) + ": " + hipGetErrorString(e));
}
#define ETN_HIP(call) ::etn::hip_check((call), #call)

inline size_t voxel_bytes(etn_voxel_type t)
{
    switch (t) {
        case ETN_VOXEL_FLOAT32: return 4;
        case ETN_VOXEL_FLOAT16: return 2;
        case ETN_VOXEL_BFLOAT16: return 2;
    }
    fail("unknown voxel type");
}

// Owning device allocation.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes) { allocate(bytes); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(std::exchange(o.ptr_, nullptr)), bytes_(std::exchange(o.bytes_, 0)) {}
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept
    {
        if (this != &o) {
            release();
            ptr_ = std::exchange(o.ptr_, nullptr);
            bytes_ = std::exchange(o.bytes_, 0);
        }
        return *this;
    }
    void allocate(size_t bytes)
    {
        release();
        if (bytes) ETN_HIP(hipMalloc(&ptr_, bytes));
        bytes_ = bytes;
    }
    void release()
    {
        if (ptr_) (void)hipFree(ptr_);
        ptr_ = nullptr;
        bytes_ = 0;
    }
    void* get() const { return ptr_; }
    template <typename T> T* as() const { return static_cast<T*>(ptr_); }
    size_t bytes() const { return bytes_; }

private:
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
};

}  // namespace etn
