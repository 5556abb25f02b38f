// device_buffer.hpp — an owned, zero-filled HIP device allocation (implementation in solver.cpp).
#pragma once

#include <cstddef>

namespace dh {

class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() { release(); }
  void allocate(size_t bytes);  // hipMalloc + zero fill; throws std::runtime_error on failure
  void adopt(void* p, size_t bytes) {
    release();
    ptr_ = p;
    bytes_ = bytes;
  }
  void release();
  template <class T>
  T* as() const {
    return static_cast<T*>(ptr_);
  }
  size_t bytes() const { return bytes_; }

 private:
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
};

}  // namespace dh
