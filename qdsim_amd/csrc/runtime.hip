// Load-time check of the HIP runtime the library is bound to.
//
// No reference counterpart.  The library links against libamdhip64; a process that
// loads it before another component brings in a second copy of the HIP runtime
// (torch ships its own) ends up with two runtimes, and kernels launched through the
// library's copy then fail ("no ROCm-capable device", "invalid device function")
// while the other copy sees the device.  qds_runtime_check makes that state an
// explicit error: it launches one trivial kernel through the library's runtime and
// reads its result back, and on failure names the libamdhip64 images mapped into
// the process.
#include "qds_common.h"

#include <link.h>

#include <cstring>
#include <string>

namespace {

__global__ void probe_kernel(int* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = 0x51d5;
}

struct RuntimeImages {
  int count = 0;
  std::string names;
};

int count_runtime(struct dl_phdr_info* info, size_t, void* data) {
  auto* r = static_cast<RuntimeImages*>(data);
  const char* name = info->dlpi_name;
  if (name && std::strstr(name, "libamdhip64")) {
    ++r->count;
    if (!r->names.empty()) r->names += ", ";
    r->names += name;
  }
  return 0;
}

}  // namespace

QDS_API int qds_runtime_check(void) {
  RuntimeImages imgs;
  dl_iterate_phdr(count_runtime, &imgs);
  const char* hint = imgs.count > 1 ? "; several HIP runtime images are mapped: load the one that owns the device "
                                      "buffers (e.g. import torch) before libqdsim_hip.so"
                                    : "";
  int n = 0;
  hipError_t rc = hipGetDeviceCount(&n);
  if (rc != hipSuccess || n <= 0) {
    qds::set_error("qds_runtime_check: the library's HIP runtime sees no device (%s; runtime images: %s)%s",
                   hipGetErrorString(rc), imgs.names.c_str(), hint);
    return QDS_EHIP;
  }
  int* d = nullptr;
  rc = hipMalloc(&d, sizeof(int));
  int h = 0;
  if (rc == hipSuccess) {
    hipLaunchKernelGGL(probe_kernel, dim3(1), dim3(64), 0, 0, d);
    rc = hipGetLastError();
    if (rc == hipSuccess) rc = hipMemcpy(&h, d, sizeof(int), hipMemcpyDeviceToHost);
    (void)hipFree(d);
  }
  if (rc != hipSuccess || h != 0x51d5) {
    qds::set_error("qds_runtime_check: a probe kernel did not run through the library's HIP runtime (%s; runtime "
                   "images: %s)%s",
                   hipGetErrorString(rc), imgs.names.c_str(), hint);
    return QDS_EHIP;
  }
  return QDS_OK;
}

// Number of HIP runtime images mapped into the process (1 expected).
QDS_API int qds_runtime_images(void) {
  RuntimeImages imgs;
  dl_iterate_phdr(count_runtime, &imgs);
  return imgs.count;
}
