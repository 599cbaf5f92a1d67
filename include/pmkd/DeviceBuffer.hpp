// DeviceBuffer.hpp -- owning device array (the CUDADataBlock<T> role:
// GetDeviceData / GetSize / Extend).
#pragma once
#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

#include <cstddef>
#include <utility>
#include <vector>

#include "Error.hpp"

namespace pmkd {

template <class T>
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t n) { Extend(n); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    DeviceBuffer(DeviceBuffer &&o) noexcept { *this = std::move(o); }
    DeviceBuffer &operator=(DeviceBuffer &&o) noexcept {
        std::swap(ptr_, o.ptr_); std::swap(size_, o.size_); std::swap(cap_, o.cap_);
        return *this;
    }
    ~DeviceBuffer() { if (ptr_) (void)hipFree(ptr_); }

    // grow to n elements, keeping the contents (never shrinks the allocation)
    void Extend(size_t n) {
        if (n > cap_) {
            T *np = nullptr;
            if (hipMalloc(&np, n * sizeof(T)) != hipSuccess) throw Exception(PM_ERR_HIP, "hipMalloc failed");
            if (ptr_) {
                if (size_ && hipMemcpy(np, ptr_, size_ * sizeof(T), hipMemcpyDeviceToDevice) != hipSuccess)
                    throw Exception(PM_ERR_HIP, "hipMemcpy failed");
                (void)hipFree(ptr_);
            }
            ptr_ = np;
            cap_ = n;
        }
        size_ = n;
    }
    void Upload(const std::vector<T> &h) {
        Extend(h.size());
        if (!h.empty() && hipMemcpy(ptr_, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            throw Exception(PM_ERR_HIP, "hipMemcpy failed");
    }
    std::vector<T> Download() const {
        std::vector<T> h(size_);
        if (size_ && hipMemcpy(h.data(), ptr_, size_ * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
            throw Exception(PM_ERR_HIP, "hipMemcpy failed");
        return h;
    }
    T *GetDeviceData() const { return ptr_; }
    size_t GetSize() const { return size_; }

private:
    T *ptr_ = nullptr;
    size_t size_ = 0, cap_ = 0;
};

}  // namespace pmkd
