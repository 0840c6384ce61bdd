// Internal plumbing shared by the HIP translation units of libhfx_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace hfx {

struct HipError : std::runtime_error {
  explicit HipError(const std::string& s) : std::runtime_error(s) {}
};

#define HFX_HIP(call)                                                                                  \
  do {                                                                                                 \
    hipError_t err_ = (call);                                                                          \
    if (err_ != hipSuccess)                                                                            \
      throw ::hfx::HipError(std::string(#call) + " failed: " + hipGetErrorString(err_) + " (" __FILE__ \
                            ":" + std::to_string(__LINE__) + ")");                                     \
  } while (0)

// Owning device buffer.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    HFX_HIP(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  void upload(const T* h, size_t count, hipStream_t s) {
    alloc(count);
    if (count) HFX_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
  void download(T* h, size_t count, hipStream_t s) const {
    if (count > n) throw std::runtime_error("DevBuf::download: count exceeds buffer");
    if (count) HFX_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
};

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// Launch facts of one kernel kept per device, thread-safe: the dynamic-LDS
// limit is raised once on each device (hipFuncSetAttribute acts on the
// current device) and the count of resident workgroups is queried once per
// device (it depends on that device's CUs).  One static instance per kernel.
struct PerDeviceLaunch {
  static constexpr int MAXD = 64;
  std::mutex mu;
  bool attr[MAXD] = {};
  int resident[MAXD] = {};

  static int current() {
    int dev = 0;
    HFX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAXD) throw std::runtime_error("kernel launch: device id out of range");
    return dev;
  }
  // raise the dynamic LDS limit of `kern` to `lds` bytes on the current device
  void lds_attr(const void* kern, size_t lds) {
    const int dev = current();
    std::lock_guard<std::mutex> lock(mu);
    if (!attr[dev]) {
      HFX_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      attr[dev] = true;
    }
  }
  // workgroups of `kern` resident at once on the current device
  int resident_wgs(const void* kern, int threads, size_t lds) {
    const int dev = current();
    std::lock_guard<std::mutex> lock(mu);
    if (resident[dev] == 0) {
      int cus = 0, per = 0;
      HFX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      HFX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, threads, lds));
      resident[dev] = (per > 0 ? per : 1) * (cus > 0 ? cus : 1);
    }
    return resident[dev];
  }
};

}  // namespace hfx
