// Internal helpers shared by the translation units of libmfbem.
#pragma once

#include <hip/hip_runtime.h>

namespace mfb {
void set_error(const char* fmt, ...);
}

struct mfb_qnet_s;
// device pointer of a network's parameters (used by the assembly's learned rule choice)
float* mfb_qnet_device_params(struct mfb_qnet_s* q);
// chosen class per row of X (device arrays) on stream s, without synchronisation
int mfb_qnet_classes_device(struct mfb_qnet_s* q, const float* X, int n, float confidence,
                            int* classes, hipStream_t s);

#define MFB_HIP(call)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) {                                                                    \
      mfb::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
      return MFB_ERR_HIP;                                                                      \
    }                                                                                          \
  } while (0)

#define MFB_ARG(cond, msg)              \
  do {                                  \
    if (!(cond)) {                      \
      mfb::set_error("%s", msg);        \
      return MFB_ERR_ARG;               \
    }                                   \
  } while (0)
