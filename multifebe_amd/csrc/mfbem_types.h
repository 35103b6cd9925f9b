// Plain-data structures shared by the host side and the kernels of libmfbem.
#pragma once

namespace mfb {

// Quasi-singular quadrature-order estimator parameters, [curve][integrand] with integrand 0 = ln r
// and 1..7 = 1/r^n (fbem_qs_parameters, lib/fbem/src/quasisingular_integration.f90:294-309).
struct QsiParams {
  double dmin[3][8], dmax[3][8], a0[3][8], a1[3][8], b1[3][8];
  double tdmin[3][8], tdmax[3][8], ta0[3][8], ta1[3][8], tb1[3][8];
};

// Complex coefficients of the harmonic fundamental solution, (re, im) pairs
// (fbem_bem_harela3d_parameters, lib/fbem/src/bem_harela3d.f90:118-134; u* and t* parts only).
struct HarParams {
  double k1[2], k2[2];
  double psi[6][2], chi[6][2];
  double T1[10][2], T2[9][2], T3[9][2];
  double cte_u[2], cte_t[2];
};

// Kernel-family constants of one assembly call.
struct KernelParams {
  int harmonic;
  double cteu1, cteu2, ctet1, ctet2;  // static Kelvin constants (bem_staela3d.f90:616-619)
  HarParams h;
};

constexpr int kMaxPrecalsets = 16;
constexpr int kPsStride = 16;  // doubles per precalculated point: x(3) n(3) phi*J*w(9) pad(1)

void host_gl_tables(double* x11, double* w11, double* x01, double* w01);  // [33][32] each
void host_qsi_params(double relative_error, QsiParams* p);
void host_har_params(double lam_re, double lam_im, double mu_re, double mu_im, double rho,
                     double omega, HarParams* p);

}  // namespace mfb
