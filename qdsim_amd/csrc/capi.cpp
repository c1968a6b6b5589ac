// Library-level entry points of the qdsim_hip C-ABI: error reporting, version and
// device queries.
#include "qds_common.h"

#include <cstring>

namespace {
thread_local char g_err[1024] = {0};
}

namespace qds {
void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
}  // namespace qds

QDS_API const char* qds_last_error(void) { return g_err; }

QDS_API int qds_version(void) { return 100; }

// Name of the device the library will launch on ("gfx950" expected) and its CU count.
QDS_API int qds_device_info(char* arch, int32_t arch_len, int32_t* cu_count) {
  int dev = 0;
  QDS_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  QDS_HIP(hipGetDeviceProperties(&prop, dev));
  if (arch && arch_len > 0) {
    std::strncpy(arch, prop.gcnArchName, arch_len - 1);
    arch[arch_len - 1] = 0;
  }
  if (cu_count) *cu_count = prop.multiProcessorCount;
  return QDS_OK;
}
