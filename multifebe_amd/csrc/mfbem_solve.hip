// Dense LU solve (placeholder until the blocked LU lands): every entry point reports
// MFB_ERR_UNSUPPORTED.
#include <hip/hip_runtime.h>

#include "mfbem.h"
#include "mfbem_internal.h"

extern "C" int mfb_dgesv_device(int, int, double*, long long, int*, double*, long long, int,
                                mfb_solve_report*, void*) {
  mfb::set_error("dense solve not built yet");
  return MFB_ERR_UNSUPPORTED;
}
extern "C" int mfb_zgesv_device(int, int, double*, long long, int*, double*, long long, int,
                                mfb_solve_report*, void*) {
  mfb::set_error("dense solve not built yet");
  return MFB_ERR_UNSUPPORTED;
}
extern "C" int mfb_solve_lse_r(int, double*, int*, int, double*, char*, double*, double*, int,
                               double*, int, int, int, int, int, mfb_solve_report*) {
  mfb::set_error("dense solve not built yet");
  return MFB_ERR_UNSUPPORTED;
}
extern "C" int mfb_solve_lse_c(int, double*, int*, int, double*, char*, double*, double*, int,
                               double*, int, int, int, int, int, mfb_solve_report*) {
  mfb::set_error("dense solve not built yet");
  return MFB_ERR_UNSUPPORTED;
}
