// Scratch of the no-allocation forward (vd_forward_device_noalloc): one
// file-scope device array, present on every GPU the library runs on until the
// process ends.  Host code never passes the array's name to a kernel or a copy
// (in host code it only names a host-side stand-in): noalloc_scratch() asks
// the runtime once per device, with that device current, for the array's
// device address and hands that out.
#include <hip/hip_runtime.h>

#include "vdhmm_internal.h"

namespace vd {
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;  // 64 MiB
constexpr int NOALLOC_MAX_DEVICES = 64;
}  // namespace vd

static __device__ __align__(256) unsigned char scratch[vd::NOALLOC_SCRATCH_BYTES];

namespace vd {

hipError_t noalloc_scratch(int dev, void** ptr, size_t* bytes) {
  static void* addr[NOALLOC_MAX_DEVICES] = {};
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
  if (!addr[dev]) {
    hipError_t e = hipSetDevice(dev);
    if (e != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[dev] = p;
  }
  *ptr = addr[dev];
  *bytes = NOALLOC_SCRATCH_BYTES;
  return hipSuccess;
}

// Parameter tables travel as kernel arguments (each below the 4 KiB argument
// limit) and are written into the scratch by one wave: no host-to-device copy.
template <typename T>
__global__ void k_put_table(T v, T* dst) {
  const uint32_t* s = reinterpret_cast<const uint32_t*>(&v);
  uint32_t* d = reinterpret_cast<uint32_t*>(dst);
  for (int i = threadIdx.x; i < (int)(sizeof(T) / 4); i += blockDim.x) d[i] = s[i];
}

struct Exps {
  int32_t e[64];
};

hipError_t launch_put_tables(const QTab<float>& t32, const QTab<double>& t64, const int32_t* E, QTab<float>* d32,
                             QTab<double>* d64, int32_t* dE, hipStream_t st) {
  static_assert(sizeof(QTab<double>) % 4 == 0 && sizeof(QTab<float>) % 4 == 0, "tables are copied in words");
  static_assert(sizeof(QTab<double>) < 4096 && sizeof(QTab<float>) < 4096, "kernel argument limit");
  Exps ex;
  for (int q = 0; q < 64; ++q) ex.e[q] = E[q];
  hipLaunchKernelGGL(k_put_table<QTab<float>>, dim3(1), dim3(64), 0, st, t32, d32);
  hipLaunchKernelGGL(k_put_table<QTab<double>>, dim3(1), dim3(64), 0, st, t64, d64);
  hipLaunchKernelGGL(k_put_table<Exps>, dim3(1), dim3(64), 0, st, ex, reinterpret_cast<Exps*>(dE));
  return hipGetLastError();
}

}  // namespace vd
