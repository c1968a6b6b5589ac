// Shared helpers for the qdsim_hip C-ABI library (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <string>

namespace qds {

// Error state of the last failing call on this host thread.
void set_error(const char* fmt, ...);

enum DType : int32_t { F32 = 1, F64 = 2, C64 = 3, C128 = 4 };

inline int dtype_size(int32_t dt) {
  switch (dt) {
    case F32: return 4;
    case F64: return 8;
    case C64: return 8;
    case C128: return 16;
    default: return 0;
  }
}

}  // namespace qds

#define QDS_OK 0
#define QDS_EINVAL (-1)
#define QDS_EHIP (-2)
#define QDS_ENOMEM (-3)
#define QDS_EWORKSPACE (-4)

#define QDS_CHECK_ARG(cond, ...)          \
  do {                                    \
    if (!(cond)) {                        \
      qds::set_error(__VA_ARGS__);        \
      return QDS_EINVAL;                  \
    }                                     \
  } while (0)

#define QDS_HIP(call)                                                         \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      qds::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),   \
                     __FILE__, __LINE__);                                     \
      return QDS_EHIP;                                                        \
    }                                                                         \
  } while (0)

#define QDS_LAUNCH_CHECK()                                                     \
  do {                                                                         \
    hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess) {                                                    \
      qds::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), \
                     __FILE__, __LINE__);                                      \
      return QDS_EHIP;                                                         \
    }                                                                          \
  } while (0)

#define QDS_API extern "C" __attribute__((visibility("default")))

static inline hipStream_t qds_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

static inline unsigned qds_blocks(int64_t work, int per_block, int64_t cap = 1 << 20) {
  int64_t b = (work + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return static_cast<unsigned>(b);
}
