// Device-side building blocks of the BEM assembly (gfx950, FP64).
//
// Everything here runs on the GPU inside the assembly kernels: quad9/line3 shape functions, the
// nearest-element-point search, the quasi-singular quadrature-order estimators, the Telles cubic
// transformation, the polar transformation of the singular case and the Kelvin / complex Stokes
// fundamental solutions.  Each function restates the reference routine named in its comment
// (paths relative to the MultiFEBE tree) so that the GPU takes the same integration decisions as
// the reference for every (collocation point, element) pair.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "mfbem_types.h"

namespace mfb {

// ---------------------------------------------------------------------------------------------
// complex double
// ---------------------------------------------------------------------------------------------
struct zd {
  double re, im;
};
__host__ __device__ inline zd mkz(double r, double i = 0.0) { return zd{r, i}; }
__host__ __device__ inline zd operator+(zd a, zd b) { return zd{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline zd operator-(zd a, zd b) { return zd{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline zd operator-(zd a) { return zd{-a.re, -a.im}; }
__host__ __device__ inline zd operator*(zd a, zd b) {
  return zd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline zd operator*(zd a, double s) { return zd{a.re * s, a.im * s}; }
__host__ __device__ inline zd operator*(double s, zd a) { return zd{s * a.re, s * a.im}; }
__host__ __device__ inline zd operator/(zd a, double s) { return zd{a.re / s, a.im / s}; }
__host__ __device__ inline zd& operator+=(zd& a, zd b) { a.re += b.re; a.im += b.im; return a; }
__host__ __device__ inline zd& operator*=(zd& a, double s) { a.re *= s; a.im *= s; return a; }

// ---------------------------------------------------------------------------------------------
// constants (lib/fbem/src/numerical.f90:74-93, shape_functions.f90:280)
// ---------------------------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846264338328;
constexpr double k2Pi = 6.28318530717958623199592693709;
constexpr double kPi_2 = 1.57079632679489661923132169164;
constexpr double kXiTol = 0.5e-12;

// Gauss-Legendre rules 1..32 on [-1,1] and [0,1] (rule n occupies row n, n entries)
struct GLTables {
  double x11[33][32], w11[33][32], x01[33][32], w01[33][32];
};

__device__ inline const GLTables& gl();  // defined by the translation unit that owns the table

// quad9 local node coordinates and edges (line3: end, end, middle)
__device__ constexpr double kXiQ9[9][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}, {0, -1},
                                           {1, 0},   {0, 1},  {-1, 0}, {0, 0}};
__device__ constexpr int kEdgeQ9[4][3] = {{0, 1, 4}, {1, 2, 5}, {2, 3, 6}, {3, 0, 7}};

// Geometry, nearest-point search, estimators and polar sub-triangles decide which rule integrates
// a pair (node ties, ceilings, 1e-12 tests).  They are compiled without FMA contraction so that
// every rounding is that of a plain FP64 evaluation of the same expression, as in the reference;
// with contraction, mirror-symmetric nodes at equal distance broke ties the other way and the
// search converged to the mirror-image minimum.  The kernel evaluation below keeps FMA.
#pragma clang fp contract(off)

// ---------------------------------------------------------------------------------------------
// shape functions (resources_shape_functions/phi_quad9.rc, dphidxi{1,2}_quad9.rc, phi_quad4.rc,
// dphidxi{1,2}_quad4.rc, phi_line3.rc, dphidxi_line3.rc), continuous elements
// ---------------------------------------------------------------------------------------------
__host__ __device__ inline void phi_q9(double x1, double x2, double* p) {
  double a3 = 0.25 * x1 * (x1 + 1.0), a4 = 0.25 * x1 * (x1 - 1.0);
  double a5 = x2 * (x2 + 1.0), a6 = x2 * (x2 - 1.0);
  double a7 = 1.0 - x1 * x1, a8 = 1.0 - x2 * x2;
  p[0] = a4 * a6; p[1] = a3 * a6; p[2] = a3 * a5; p[3] = a4 * a5;
  p[4] = 0.5 * a6 * a7; p[5] = 2.0 * a3 * a8; p[6] = 0.5 * a5 * a7; p[7] = 2.0 * a4 * a8;
  p[8] = a7 * a8;
}

__device__ inline void dphi_q9(double x1, double x2, double* d1, double* d2) {
  {
    double a3 = 2.0 * x1 + 1.0, a4 = 2.0 * x1 - 1.0, a5 = x2 + 1.0, a6 = x2 - 1.0;
    double a7 = 0.25 * x2, a8 = a5 * a6, a9 = -0.5 * a8, a10 = -x1 * x2;
    d1[0] = a7 * a4 * a6; d1[1] = a7 * a3 * a6; d1[2] = a7 * a3 * a5; d1[3] = a7 * a4 * a5;
    d1[4] = a10 * a6; d1[5] = a9 * a3; d1[6] = a10 * a5; d1[7] = a9 * a4; d1[8] = 2.0 * x1 * a8;
  }
  {
    double a3 = 2.0 * x2 + 1.0, a4 = 2.0 * x2 - 1.0, a5 = x1 + 1.0, a6 = x1 - 1.0;
    double a7 = 0.25 * x1, a8 = a5 * a6, a9 = -0.5 * a8, a10 = -x1 * x2;
    d2[0] = a7 * a6 * a4; d2[1] = a7 * a5 * a4; d2[2] = a7 * a5 * a3; d2[3] = a7 * a6 * a3;
    d2[4] = a9 * a4; d2[5] = a10 * a5; d2[6] = a9 * a3; d2[7] = a10 * a6; d2[8] = 2.0 * x2 * a8;
  }
}

__device__ inline void phi_l3(double xi, double* p) {
  double a2 = 0.5 * xi, a3 = xi - 1.0, a4 = xi + 1.0;
  p[0] = a2 * a3; p[1] = a2 * a4; p[2] = -a3 * a4;
}

__device__ inline void dphi_l3(double xi, double* d) {
  d[0] = xi - 0.5; d[1] = xi + 0.5; d[2] = -2.0 * xi;
}

__device__ inline double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ inline void cross3(const double* a, const double* b, double* n) {
  n[0] = a[1] * b[2] - a[2] * b[1];
  n[1] = a[2] * b[0] - a[0] * b[2];
  n[2] = a[0] * b[1] - a[1] * b[0];
}

// x, T1 = dx/dxi1, T2 = dx/dxi2 of a quad9 (X node-major, 27 doubles)
__device__ inline void q9_geom(const double* X, double x1, double x2, double* x, double* T1,
                               double* T2) {
  double p[9], d1[9], d2[9];
  phi_q9(x1, x2, p);
  dphi_q9(x1, x2, d1, d2);
#pragma unroll
  for (int c = 0; c < 3; c++) { x[c] = 0; T1[c] = 0; T2[c] = 0; }
#pragma unroll
  for (int k = 0; k < 9; k++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      x[c] += p[k] * X[3 * k + c];
      T1[c] += d1[k] * X[3 * k + c];
      T2[c] += d2[k] * X[3 * k + c];
    }
}

// Double-double arithmetic (TwoSum; TwoProd through fma): the residual x(xi) - x_i and the tangents
// of the nearest-point iterations to about twice FP64 precision.  The reference iterates in real128.
// In FP64, far outside an element (|xi| >> 1) the quadratic shape functions are large and cancel in
// the residual, and a point off the element's plane leaks its normal distance into the step through
// the rounding of the tangents' normal components: both leave a step noise of 1e-14..1e-13 that keeps
// the 1e-14 convergence test from passing where the real128 iteration converges.
struct ddv {
  double hi, lo;
};
__host__ __device__ inline ddv dd_two_sum(double a, double b) {
  const double s = a + b, v = s - a;
  return {s, (a - (s - v)) + (b - v)};
}
__host__ __device__ inline ddv dd_norm(double h, double l) {
  const double s = h + l;
  return {s, l - (s - h)};
}
__host__ __device__ inline ddv dd_add(ddv a, ddv b) {
  const ddv s = dd_two_sum(a.hi, b.hi);
  return dd_norm(s.hi, s.lo + a.lo + b.lo);
}
__host__ __device__ inline ddv dd_mul(ddv a, ddv b) {
  const double p = a.hi * b.hi;
  return dd_norm(p, fma(a.hi, b.hi, -p) + (a.hi * b.lo + a.lo * b.hi));
}
__host__ __device__ inline ddv dd_mul_d(ddv a, double b) {
  const double p = a.hi * b;
  return dd_norm(p, fma(a.hi, b, -p) + a.lo * b);
}
// 1D quadratic Lagrange functions L and derivatives D at t for the nodes -1, 0, 1
__host__ __device__ inline void dd_lagrange3(double t, ddv* L, ddv* D) {
  const ddv tm = dd_two_sum(t, -1.0), tp = dd_two_sum(t, 1.0), om = dd_two_sum(1.0, -t);
  const ddv a = dd_mul_d(tm, t), c = dd_mul_d(tp, t);
  L[0] = {0.5 * a.hi, 0.5 * a.lo};
  L[1] = dd_mul(om, tp);
  L[2] = {0.5 * c.hi, 0.5 * c.lo};
  D[0] = dd_two_sum(t, -0.5);
  D[1] = {-2.0 * t, 0.0};
  D[2] = dd_two_sum(t, 0.5);
}
// x(xi), dx/dxi1, dx/dxi2 of a quad9 (node k at kXiQ9[k]), each rounded once from double-double
__host__ __device__ inline void q9_geom_dd(const double* X, double x1, double x2, double* x,
                                           double* T1, double* T2) {
  const int sa[9] = {0, 2, 2, 0, 1, 2, 1, 0, 1}, sb[9] = {0, 0, 2, 2, 0, 1, 2, 1, 1};
  ddv La[3], Da[3], Lb[3], Db[3];
  ddv ax[3] = {{0, 0}, {0, 0}, {0, 0}}, a1[3] = {{0, 0}, {0, 0}, {0, 0}}, a2[3] = {{0, 0}, {0, 0}, {0, 0}};
  dd_lagrange3(x1, La, Da);
  dd_lagrange3(x2, Lb, Db);
  for (int k = 0; k < 9; k++) {
    const ddv ph = dd_mul(La[sa[k]], Lb[sb[k]]);
    const ddv d1 = dd_mul(Da[sa[k]], Lb[sb[k]]), d2 = dd_mul(La[sa[k]], Db[sb[k]]);
    for (int c = 0; c < 3; c++) {
      ax[c] = dd_add(ax[c], dd_mul_d(ph, X[3 * k + c]));
      a1[c] = dd_add(a1[c], dd_mul_d(d1, X[3 * k + c]));
      a2[c] = dd_add(a2[c], dd_mul_d(d2, X[3 * k + c]));
    }
  }
  for (int c = 0; c < 3; c++) {
    x[c] = ax[c].hi + ax[c].lo;
    T1[c] = a1[c].hi + a1[c].lo;
    T2[c] = a2[c].hi + a2[c].lo;
  }
}
// x(xi) and dx/dxi of a line3 (nodes at -1, 1, 0), rounded once from double-double
__host__ __device__ inline void l3_geom_dd(const double* X, double t, double* x, double* T) {
  ddv L[3], D[3], ax[3] = {{0, 0}, {0, 0}, {0, 0}}, at[3] = {{0, 0}, {0, 0}, {0, 0}};
  dd_lagrange3(t, L, D);
  const int pos[3] = {0, 2, 1};
  for (int m = 0; m < 3; m++)
    for (int c = 0; c < 3; c++) {
      ax[c] = dd_add(ax[c], dd_mul_d(L[pos[m]], X[3 * m + c]));
      at[c] = dd_add(at[c], dd_mul_d(D[pos[m]], X[3 * m + c]));
    }
  for (int c = 0; c < 3; c++) {
    x[c] = ax[c].hi + ax[c].lo;
    T[c] = at[c].hi + at[c].lo;
  }
}

__host__ __device__ inline void q9_x(const double* X, double x1, double x2, double* x) {
  double p[9];
  phi_q9(x1, x2, p);
  for (int c = 0; c < 3; c++) x[c] = 0;
  for (int k = 0; k < 9; k++)
    for (int c = 0; c < 3; c++) x[c] += p[k] * X[3 * k + c];
}

// ---------------------------------------------------------------------------------------------
// line lengths (geometry.f90:1469-1504, 2794-2882) and characteristic length (:3357-3390)
// ---------------------------------------------------------------------------------------------
__device__ inline double jac_l3(const double* X, double xi) {
  double d[3], t[3] = {0, 0, 0};
  dphi_l3(xi, d);
  for (int k = 0; k < 3; k++)
    for (int c = 0; c < 3; c++) t[c] += d[k] * X[3 * k + c];
  return sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
}

__device__ inline double length_l3(const double* X, double tol) {
  const GLTables& G = gl();
  const int step = 3;
  double length = jac_l3(X, G.x11[1][0]) * G.w11[1][0], err = tol + 1.0;
  int j = 1 + step;
  while (err >= tol) {
    double old = length;
    length = 0;
    for (int i = 0; i < j; i++) length += jac_l3(X, G.x11[j][i]) * G.w11[j][i];
    err = fabs((old - length) / length);
    j += step;
    if (j >= 32) {
      length = 0;
      for (int i = 0; i < 32; i++) length += jac_l3(X, G.x11[32][i]) * G.w11[32][i];
      break;
    }
  }
  return length;
}

__device__ inline double clen_q9(const double* X, double tol) {
  double l = 0;
  for (int e = 0; e < 4; e++) {
    double xe[9];
    for (int k = 0; k < 3; k++)
      for (int c = 0; c < 3; c++) xe[3 * k + c] = X[3 * kEdgeQ9[e][k] + c];
    double le = length_l3(xe, tol);
    if (e == 0 || le > l) l = le;
  }
  return l;
}

// ---------------------------------------------------------------------------------------------
// quasi-singular rule estimators (quasisingular_integration.f90:402-715); the parameter table is
// filled on the host (mfbem_qsi.cpp) exactly as fbem_qs_calculate_parameters does
// ---------------------------------------------------------------------------------------------
__device__ inline int qs_estimate(const QsiParams& p, int telles, int ndim, int f, double d,
                                  const double* barxi) {
  const double(*dmin)[8] = telles ? p.tdmin : p.dmin;
  const double(*dmax)[8] = telles ? p.tdmax : p.dmax;
  const double(*a0)[8] = telles ? p.ta0 : p.a0;
  const double(*a1)[8] = telles ? p.ta1 : p.a1;
  const double(*b1)[8] = telles ? p.tb1 : p.b1;
  int n;
  if (d <= 2.0) {
    double ld = log10(d), N1, N2;
    if (ndim == 1 && !(fabs(barxi[0]) < 1.0)) {
      if (d <= dmin[1][f]) return 0;
      N2 = d >= dmax[1][f] ? 2.0 : pow(10.0, (a0[1][f] + a1[1][f] * ld) / (1.0 + b1[1][f] * ld));
      n = (int)ceil(N2);
    } else {
      if (d <= dmin[0][f]) return 0;
      N1 = d >= dmax[0][f] ? 2.0 : pow(10.0, (a0[0][f] + a1[0][f] * ld) / (1.0 + b1[0][f] * ld));
      if (d <= dmin[1][f]) return 0;
      N2 = d >= dmax[1][f] ? 2.0 : pow(10.0, (a0[1][f] + a1[1][f] * ld) / (1.0 + b1[1][f] * ld));
      if (ndim == 1)
        n = (int)ceil((barxi[0] * barxi[0]) * (N2 - N1) + N1);
      else
        n = (int)ceil((barxi[0] * barxi[0]) * (barxi[1] * barxi[1]) * (N2 - N1) + N1);
    }
  } else {
    if (d >= p.dmax[2][f]) {
      n = 2;
    } else {
      double ld = log10(d);
      n = (int)round(pow(10.0, (p.a0[2][f] + p.a1[2][f] * ld) / (1.0 + p.b1[2][f] * ld)));
    }
  }
  return n > 30 ? 0 : n;
}

// ---------------------------------------------------------------------------------------------
// Telles transformation (telles_transformation.f90:77-232)
// ---------------------------------------------------------------------------------------------
__device__ inline double telles_barr(double d) {
  double r = d < 3.0 ? d / (0.89039 * d + 0.32883) : 1.0;
  return r > 1.0 ? 1.0 : r;
}

__device__ inline void telles11(double bxi, double br, double* c) {
  double w = bxi / (1.0 + 2.0 * br);
  double p = 1.0 / (3.0 * (1.0 + 2.0 * br)) * (3.0 - 2.0 * br - 3.0 * w * bxi);
  double q = w / 2.0 * ((3.0 - 2.0 * br) / (1.0 + 2.0 * br) - 2.0 * (w * w) - 1.0);
  double R2 = sqrt(q * q + p * p * p);
  double R31 = -q + R2, R32 = -q - R2;
  const double d13 = 1.0 / 3.0;
  R31 = R31 >= 0 ? pow(R31, d13) : -pow(fabs(R31), d13);
  R32 = R32 >= 0 ? pow(R32, d13) : -pow(fabs(R32), d13);
  double g = R31 + R32 + w;
  double Q = 1.0 + 3.0 * (g * g);
  c[0] = (1.0 - br) / Q;
  c[1] = -3.0 * g * c[0];
  c[2] = (br + 3.0 * (g * g)) / Q;
  c[3] = -c[1];
}

__device__ inline void telles_xj(const double* c, double g, double* xi, double* j) {
  *xi = c[0] * (g * g * g) + c[1] * (g * g) + c[2] * g + c[3];
  *j = 3.0 * c[0] * (g * g) + 2.0 * c[1] * g + c[2];
}

// ---------------------------------------------------------------------------------------------
// nearest element point (geometry.f90:4263-4316, 4329-4440, 4574-4787, 5072-5408, 5519-5572).
// The reference runs the Newton steps in real128; here they run in FP64 (decision agreement with
// the oracle is checked by the parity tests).
// ---------------------------------------------------------------------------------------------
__device__ inline bool check_xi_q(const double* xi) {
  return !(xi[0] < -1.0 - kXiTol || xi[0] > 1.0 + kXiTol || xi[1] < -1.0 - kXiTol ||
           xi[1] > 1.0 + kXiTol);
}

__device__ inline double dist_l3(const double* X, const double* p, double xi) {
  double ph[3], x[3] = {0, 0, 0};
  phi_l3(xi, ph);
  for (int k = 0; k < 3; k++)
    for (int c = 0; c < 3; c++) x[c] += ph[k] * X[3 * k + c];
  double v[3] = {x[0] - p[0], x[1] - p[1], x[2] - p[2]};
  return sqrt(dot3(v, v));
}

__device__ inline void sample_l3(const double* X, const double* p, double* xi_out,
                                 double* r_out) {
  const int ns = 25, nr = 14;
  double xmin = 0, rmin = 0;
  for (int k = 0; k <= ns + 1; k++) {
    double xi = -1.0 + 2.0 / (double)(ns + 1) * (double)k;
    double r = dist_l3(X, p, xi);
    if (k == 0 || r < rmin) { xmin = xi; rmin = r; }
  }
  double xold = xmin, width = 2.0 / (double)(ns + 1);
  for (int l = 0; l < 2; l++) {
    for (int k = 1; k <= nr; k++) {
      double xi = xold - width + (2.0 * width) / (double)(nr + 1) * (double)k;
      if (xi > -1.0 && xi < 1.0) {
        double r = dist_l3(X, p, xi);
        if (r < rmin) { xmin = xi; rmin = r; }
      }
    }
    xold = xmin;
    width = 2.0 * width / (double)(nr + 1);
  }
  *xi_out = xmin;
  *r_out = rmin;
}

__device__ inline double dist_q9(const double* X, const double* p, const double* xi) {
  double x[3];
  q9_x(X, xi[0], xi[1], x);
  double v[3] = {x[0] - p[0], x[1] - p[1], x[2] - p[2]};
  return sqrt(dot3(v, v));
}

__device__ inline void sample_q9(const double* X, const double* p, double* xi_out,
                                 double* r_out) {
  const int ns = 25, nr = 14;
  double xm[2] = {0, 0}, rmin = 0;
  for (int k1 = 0; k1 <= ns + 1; k1++) {
    double xi[2];
    xi[0] = -1.0 + 2.0 * (double)k1 / (double)(ns + 1);
    for (int k2 = 0; k2 <= ns + 1; k2++) {
      xi[1] = -1.0 + 2.0 * (double)k2 / (double)(ns + 1);
      double r = dist_q9(X, p, xi);
      if ((k1 == 0 && k2 == 0) || r < rmin) { xm[0] = xi[0]; xm[1] = xi[1]; rmin = r; }
    }
  }
  double xo[2] = {xm[0], xm[1]}, width = 2.0 / (double)(ns + 1);
  for (int l = 0; l < 2; l++) {
    for (int k1 = 1; k1 <= nr; k1++) {
      double xi[2];
      xi[0] = xo[0] + 2.0 * ((double)k1 - (double)(nr + 1) / 2.0) / (double)(nr + 1) * width;
      if (xi[0] > -1.0 && xi[0] < 1.0) {
        for (int k2 = 1; k2 <= nr; k2++) {
          xi[1] = xo[1] + 2.0 * ((double)k2 - (double)(nr + 1) / 2.0) / (double)(nr + 1) * width;
          if (xi[1] > -1.0 && xi[1] < 1.0) {
            double r = dist_q9(X, p, xi);
            if (r < rmin) { xm[0] = xi[0]; xm[1] = xi[1]; rmin = r; }
          }
        }
      }
    }
    xo[0] = xm[0];
    xo[1] = xm[1];
    width = width / (double)(nr + 1);
    width = 2.0 * width / (double)(nr + 1);
  }
  xi_out[0] = xm[0];
  xi_out[1] = xm[1];
  *r_out = rmin;
}

__device__ inline void nearest_l3_local(const double* X, double cl, const double* x_i,
                                        double* barxi, double* rmin, double* d);

// Nearest point of a line3 (geometry.f90:5519-5572 with :5160-5286): the nodes are recentred on
// x_i first, for the same reason as in minimize_q9.
__device__ inline void nearest_l3(const double* X, double cl, const double* x_i, double* barxi,
                                  double* rmin, double* d) {
  double Y[9];
  for (int k = 0; k < 3; k++)
    for (int c = 0; c < 3; c++) Y[3 * k + c] = X[3 * k + c] - x_i[c];
  const double zero[3] = {0.0, 0.0, 0.0};
  nearest_l3_local(Y, cl, zero, barxi, rmin, d);
}

__device__ inline void nearest_l3_local(const double* X, double cl, const double* x_i,
                                        double* barxi, double* rmin, double* d) {
  // nearest node
  const double xin[3] = {-1.0, 1.0, 0.0};
  int node = 0;
  {
    double v[3] = {X[0] - x_i[0], X[1] - x_i[1], X[2] - x_i[2]};
    *rmin = sqrt(dot3(v, v));
    for (int k = 1; k < 3; k++) {
      double w[3] = {X[3 * k] - x_i[0], X[3 * k + 1] - x_i[1], X[3 * k + 2] - x_i[2]};
      double t = sqrt(dot3(w, w));
      if (t < *rmin) { *rmin = t; node = k; }
    }
  }
  *barxi = xin[node];
  *d = *rmin / cl;
  if (*d >= 1.0) return;
  double prev = (*barxi < -1.0 || *barxi > 1.0) ? 0.0 : *barxi;
  double err_m2 = 0, err_m1 = 0;
  int info = 0;
  for (int k = 1; info == 0; k++) {
    double xb[3], db[3];
    l3_geom_dd(X, prev, xb, db);
    // the step as a correction of prev (see minimize_q9)
    double gb = 0, bb = 0;
    for (int c = 0; c < 3; c++) {
      gb += (xb[c] - x_i[c]) * db[c];
      bb += db[c] * db[c];
    }
    double step = -gb / bb;
    double next = prev + step;
    double err = fabs(step) * 0.5;
    if (err <= 1e-14) { *barxi = next; info = 1; }
    else if (k == 20) { *barxi = next; info = 2; }
    else if (k > 5 && err > err_m2) { *barxi = next; info = 3; }
    err_m2 = err_m1;
    err_m1 = err;
    prev = next;
  }
  if (info == 1) {
    if (*barxi < -1.0 || *barxi > 1.0) {
      *barxi = -1.0;
      double v0[3] = {X[0] - x_i[0], X[1] - x_i[1], X[2] - x_i[2]};
      double v1[3] = {X[3] - x_i[0], X[4] - x_i[1], X[5] - x_i[2]};
      if (sqrt(dot3(v1, v1)) < sqrt(dot3(v0, v0))) *barxi = 1.0;
    }
    *rmin = dist_l3(X, x_i, *barxi);
  } else {
    sample_l3(X, x_i, barxi, rmin);
  }
  *d = *rmin / cl;
}

// Newton minimisation of |x(xi) - x_i| (geometry.f90:5289-5408).  The reference carries the
// iteration in real128; here the element is first recentred on x_i (Y_k = X_k - x_i, exact to one
// rounding of the small differences) so that FP64 reaches the same 1e-14 convergence test even on
// small sub-elements far from the coordinate origin.
__device__ inline int minimize_q9(const double* X, const double* x_i, double* barxi,
                                  double* rmin) {
  double Y[27];
  for (int k = 0; k < 9; k++)
    for (int c = 0; c < 3; c++) Y[3 * k + c] = X[3 * k + c] - x_i[c];
  const double zero[3] = {0.0, 0.0, 0.0};
  double prev[2];
  if (!check_xi_q(barxi)) { prev[0] = 0; prev[1] = 0; }
  else { prev[0] = barxi[0]; prev[1] = barxi[1]; }
  double err_m2 = 0, err_m1 = 0;
  int info = 0;
  for (int k = 1; info == 0; k++) {
    double xb[3], t1[3], t2[3];
    q9_geom_dd(Y, prev[0], prev[1], xb, t1, t2);
    // Gauss-Newton step as a correction of prev (xb = x(prev) - x_i): the same iterate as solving
    // for the new xi directly, without the rounding of |xi| |dx/dxi| that keeps a minimiser far
    // outside the element (|xi| >> 1) jittering above the 1e-14 test in FP64
    double bb = 0, cc = 0, bc = 0, g1 = 0, g2 = 0;
    for (int c = 0; c < 3; c++) {
      bb += t1[c] * t1[c];
      cc += t2[c] * t2[c];
      bc += t1[c] * t2[c];
      g1 += xb[c] * t1[c];
      g2 += xb[c] * t2[c];
    }
    double det = bb * cc - bc * bc;
    double e0 = -(g1 * cc - bc * g2) / det, e1 = -(bb * g2 - g1 * bc) / det;
    double n0 = prev[0] + e0, n1 = prev[1] + e1;
    double err = sqrt(e0 * e0 + e1 * e1) * 0.5;
    if (err <= 1e-14) info = 1;
    else if (k == 20) info = 2;
    else if (k > 5 && err >= err_m2) info = 3;
    barxi[0] = n0;
    barxi[1] = n1;
    err_m2 = err_m1;
    err_m1 = err;
    prev[0] = n0;
    prev[1] = n1;
  }
  if (info == 1) {
    if (!check_xi_q(barxi)) {
      double best_r = 0, best_x = 0;
      int kb = 0;
      for (int ke = 0; ke < 4; ke++) {
        double ye[9];
        for (int j = 0; j < 3; j++)
          for (int c = 0; c < 3; c++) ye[3 * j + c] = Y[3 * kEdgeQ9[ke][j] + c];
        double cl = length_l3(ye, 1e-12), bx, rr, dd;
        nearest_l3(ye, cl, zero, &bx, &rr, &dd);
        if (ke == 0 || rr < best_r) { best_r = rr; best_x = bx; kb = ke; }
      }
      double ph[3];
      phi_l3(best_x, ph);
      barxi[0] = 0;
      barxi[1] = 0;
      for (int j = 0; j < 3; j++) {
        barxi[0] += ph[j] * kXiQ9[kEdgeQ9[kb][j]][0];
        barxi[1] += ph[j] * kXiQ9[kEdgeQ9[kb][j]][1];
      }
    }
    *rmin = dist_q9(Y, zero, barxi);
  }
  return info;
}

__device__ inline void nearest_q9(const double* X, double cl, const double* x_i, double* barxi,
                                  double* rmin, double* d) {
  int node = 0;
  {
    double v[3] = {X[0] - x_i[0], X[1] - x_i[1], X[2] - x_i[2]};
    *rmin = sqrt(dot3(v, v));
    for (int k = 1; k < 9; k++) {
      double w[3] = {X[3 * k] - x_i[0], X[3 * k + 1] - x_i[1], X[3 * k + 2] - x_i[2]};
      double t = sqrt(dot3(w, w));
      if (t < *rmin) { *rmin = t; node = k; }
    }
  }
  barxi[0] = kXiQ9[node][0];
  barxi[1] = kXiQ9[node][1];
  *d = *rmin / cl;
  // x_i on an element node: the node is the exact minimiser and the reference's real128 Newton
  // iteration stays on it to far below FP64 resolution, so its result is the node's xi exactly
  if (*rmin == 0.0) return;
  if (*d < 1.0) {
    if (minimize_q9(X, x_i, barxi, rmin) != 1) sample_q9(X, x_i, barxi, rmin);
    *d = *rmin / cl;
  }
}

// fbem_obtain_element_subdivision_coordinates (geometry.f90:2591-2697), quad9 and line3
__device__ inline void subdiv_q9(const double* X, const double* xs /*4x2*/, double* Xs) {
  for (int k = 0; k < 9; k++) {
    double a = kXiQ9[k][0], b = kXiQ9[k][1];
    double ps[4] = {0.25 * (1.0 - a) * (1.0 - b), 0.25 * (1.0 + a) * (1.0 - b),
                    0.25 * (1.0 + a) * (1.0 + b), 0.25 * (1.0 - a) * (1.0 + b)};
    double xi0 = 0, xi1 = 0;
    for (int i = 0; i < 4; i++) {
      xi0 += ps[i] * xs[2 * i];
      xi1 += ps[i] * xs[2 * i + 1];
    }
    q9_x(X, xi0, xi1, Xs + 3 * k);
  }
}

__device__ inline void subdiv_l3(const double* X, const double* xs, double* Xs) {
  const double xin[3] = {-1.0, 1.0, 0.0};
  for (int k = 0; k < 3; k++) {
    double a1 = 0.5 * xin[k];
    double xi = (0.5 - a1) * xs[0];
    xi += (0.5 + a1) * xs[1];
    double ph[3];
    phi_l3(xi, ph);
    for (int c = 0; c < 3; c++) {
      Xs[3 * k + c] = 0;
      for (int i = 0; i < 3; i++) Xs[3 * k + c] += ph[i] * X[3 * i + c];
    }
  }
}

#pragma clang fp contract(fast)

// ---------------------------------------------------------------------------------------------
// E_0..E_6 of exp(z) (numerical.f90:1258-1330); only E_2..E_5 are consumed by the kernels
// ---------------------------------------------------------------------------------------------
__device__ inline void zexp_dec(zd z, zd* E2, zd* E3, zd* E4, zd* E5) {
  double az = sqrt(z.re * z.re + z.im * z.im);   // |z| <~ 1e3 here: no overflow to guard
  if (az <= 1.0) {
    // E6 = z^6/6! (1 + z/7 (1 + z/8 (... (1 + z/21)))) by Horner: the smallest terms first, as
    // the reference sums its terms (downwards from k = n + 6, n = round(10^(1.176 + 0.175
    // log10|z|)) <= 15 on |z| <= 1, numerical.f90:1276); the terms it leaves out are below the
    // last bit of E6 (|z|^22/22! < 1e-21 |z|^6/6! here), so every lane takes the same 15 steps
    // (no divergent loop, compile-time 1/k)
    zd h = mkz(1.0);
#pragma unroll
    for (int k = 21; k >= 7; k--) h = mkz(1.0) + h * (z * (1.0 / k));
    const zd z2 = z * z, z3 = z2 * z, z4 = z3 * z, z5 = z4 * z, z6 = z5 * z;
    *E5 = z6 * (1.0 / 720.0) * h + z5 * (1.0 / 120.0);
    *E4 = *E5 + z4 * (1.0 / 24.0);
    *E3 = *E4 + z3 * (1.0 / 6.0);
    *E2 = *E3 + z2 * 0.5;
  } else {
    double ex = exp(z.re), s, c;
    sincos(z.im, &s, &c);
    zd ez = mkz(ex * c, ex * s);
    zd z2 = z * z, z3 = z2 * z, z4 = z3 * z;
    zd e1 = ez - mkz(1.0);
    *E2 = e1 - z;
    *E3 = *E2 - 0.5 * z2;
    *E4 = *E3 - (1.0 / 6.0) * z3;
    *E5 = *E4 - (1.0 / 24.0) * z4;
  }
}

// harmonic scalar factors at distance r (bem_harela3d.f90:660-682); full=false drops the 1/r^2
// leading terms of TT1..TT3 (regular parts used by the singular integration)
__device__ inline void har_factors(const HarParams& k, double r, bool full, zd& psi, zd& chi,
                                   zd& TT1, zd& TT2, zd& TT3) {
  double d1 = 1.0 / r, d2 = d1 * d1, d3 = d2 * d1, d4 = d3 * d1;
  auto C = [](const double* v) { return mkz(v[0], v[1]); };
  zd t1, t2, t3;
  // z = -i k r; the terms of the first wave number are summed before the second one's E_n are
  // formed (fewer values live at once)
  {
    const zd z1 = mkz(-(-k.k1[1] * r), -(k.k1[0] * r));
    zd a2, a3, a4, a5;
    zexp_dec(z1, &a2, &a3, &a4, &a5);
    a2 *= d1; a3 *= d2; a4 *= d3; a5 *= d4;
    psi = C(k.psi[0]) * d1 + C(k.psi[1]) + C(k.psi[2]) * a3 + C(k.psi[4]) * a4;
    chi = C(k.chi[0]) * d1 + C(k.chi[1]) * a2 + C(k.chi[2]) * a3 + C(k.chi[4]) * a4;
    t1 = C(k.T1[1]) + C(k.T1[2]) * a2 + C(k.T1[4]) * a3 + C(k.T1[6]) * a4 + C(k.T1[8]) * a5;
    t2 = C(k.T2[1]) + C(k.T2[3]) * a3 + C(k.T2[5]) * a4 + C(k.T2[7]) * a5;
    t3 = C(k.T3[1]) + C(k.T3[2]) * a2 + C(k.T3[3]) * a3 + C(k.T3[5]) * a4 + C(k.T3[7]) * a5;
  }
  {
    const zd z2 = mkz(-(-k.k2[1] * r), -(k.k2[0] * r));
    zd b2, b3, b4, b5;
    zexp_dec(z2, &b2, &b3, &b4, &b5);
    b2 *= d1; b3 *= d2; b4 *= d3; b5 *= d4;
    psi = psi + b2 + C(k.psi[3]) * b3 + C(k.psi[5]) * b4;
    chi = chi + b2 + C(k.chi[3]) * b3 + C(k.chi[5]) * b4;
    t1 = t1 + C(k.T1[3]) * b2 + C(k.T1[5]) * b3 + C(k.T1[7]) * b4 + C(k.T1[9]) * b5;
    t2 = t2 + C(k.T2[2]) * b2 + C(k.T2[4]) * b3 + C(k.T2[6]) * b4 + C(k.T2[8]) * b5;
    t3 = t3 + C(k.T3[4]) * b3 + C(k.T3[6]) * b4 + C(k.T3[8]) * b5;
  }
  if (full) {
    TT1 = C(k.T1[0]) * d2 + t1;
    TT2 = C(k.T2[0]) * d2 + t2;
    TT3 = C(k.T3[0]) * d2 + t3;
  } else {
    TT1 = t1;
    TT2 = t2;
    TT3 = t3;
  }
}

#pragma clang fp contract(off)

// ---------------------------------------------------------------------------------------------
// polar transformation for quadrilaterals (polar_transformation.f90:251-546)
// ---------------------------------------------------------------------------------------------
__host__ __device__ inline bool check_xi_edge_q(const double* xi) {
  if (xi[0] < -1.0 - kXiTol || xi[0] > 1.0 + kXiTol || xi[1] < -1.0 - kXiTol ||
      xi[1] > 1.0 + kXiTol)
    return false;
  return xi[0] < -1.0 + kXiTol || xi[0] > 1.0 - kXiTol || xi[1] < -1.0 + kXiTol ||
         xi[1] > 1.0 - kXiTol;
}

__host__ __device__ inline int polar_setup_q(const double* xi, int* st, double* th /*8x2*/,
                                             double* thp /*8x2*/) {
  int n = 0;
  const double t = kXiTol;
  if (!check_xi_edge_q(xi)) {
    n = 8;
    for (int i = 0; i < 8; i++) st[i] = i + 1;
  } else {
    if (xi[0] <= -1.0 + t && xi[1] <= -1.0 + t) { n = 2; st[0] = 4; st[1] = 5; }
    if (xi[0] >= 1.0 - t && xi[1] <= -1.0 + t) { n = 2; st[0] = 6; st[1] = 7; }
    if (xi[0] >= 1.0 - t && xi[1] >= 1.0 - t) { n = 2; st[0] = 8; st[1] = 1; }
    if (xi[0] <= -1.0 + t && xi[1] >= 1.0 - t) { n = 2; st[0] = 2; st[1] = 3; }
    if (n == 0) {
      if (xi[1] <= -1.0 + t) { n = 4; st[0] = 4; st[1] = 5; st[2] = 6; st[3] = 7; }
      if (xi[0] >= 1.0 - t) { n = 4; st[0] = 6; st[1] = 7; st[2] = 8; st[3] = 1; }
      if (xi[1] >= 1.0 - t) { n = 4; st[0] = 8; st[1] = 1; st[2] = 2; st[3] = 3; }
      if (xi[0] <= -1.0 + t) { n = 4; st[0] = 2; st[1] = 3; st[2] = 4; st[3] = 5; }
    }
  }
  const double x0 = xi[0], x1 = xi[1];
  for (int k = 0; k < n; k++) {
    double a, b, s0 = 0, s1 = 0;  // s: scale factor of log(tan(.)) and the shift
    double th1 = kPi - asin((-1.0 - x1) / sqrt((-1.0 - x0) * (-1.0 - x0) + (-1.0 - x1) * (-1.0 - x1)));
    double th2 = k2Pi + asin((-1.0 - x1) / sqrt((1.0 - x0) * (1.0 - x0) + (-1.0 - x1) * (-1.0 - x1)));
    double th3 = asin((1.0 - x1) / sqrt((1.0 - x0) * (1.0 - x0) + (1.0 - x1) * (1.0 - x1)));
    double th4 = kPi - asin((1.0 - x1) / sqrt((-1.0 - x0) * (-1.0 - x0) + (1.0 - x1) * (1.0 - x1)));
    switch (st[k]) {
      case 1: a = th1; b = 1.5 * kPi; s0 = 1.0 + x1; s1 = -kPi; break;
      case 2: a = 1.5 * kPi; b = th2; s0 = 1.0 + x1; s1 = -kPi; break;
      case 3: a = th2; b = k2Pi; s0 = 1.0 - x0; s1 = kPi_2; break;
      case 4: a = 0.0; b = th3; s0 = 1.0 - x0; s1 = kPi_2; break;
      case 5: a = th3; b = kPi_2; s0 = 1.0 - x1; s1 = 0.0; break;
      case 6: a = kPi_2; b = th4; s0 = 1.0 - x1; s1 = 0.0; break;
      case 7: a = th4; b = kPi; s0 = 1.0 + x0; s1 = -kPi_2; break;
      default: a = kPi; b = th1; s0 = 1.0 + x0; s1 = -kPi_2; break;
    }
    th[2 * k] = a;
    th[2 * k + 1] = b;
    if (s1 == 0.0) {
      thp[2 * k] = s0 * log(tan(0.5 * a));
      thp[2 * k + 1] = s0 * log(tan(0.5 * b));
    } else {
      thp[2 * k] = s0 * log(tan(0.5 * (a + s1)));
      thp[2 * k + 1] = s0 * log(tan(0.5 * (b + s1)));
    }
  }
  return n;
}

__device__ inline void polar_angular_q(const double* xi, int st, double thp, double* th,
                                       double* rho) {
  switch (st) {
    case 1: case 2:
      *th = 2.0 * atan(exp(thp / (1.0 + xi[1]))) + kPi;
      *rho = (-1.0 - xi[1]) / sin(*th);
      break;
    case 3: case 4:
      *th = 2.0 * atan(exp(thp / (1.0 - xi[0]))) - kPi_2;
      *rho = (1.0 - xi[0]) / cos(*th);
      break;
    case 5: case 6:
      *th = 2.0 * atan(exp(thp / (1.0 - xi[1])));
      *rho = (1.0 - xi[1]) / sin(*th);
      break;
    default:
      *th = 2.0 * atan(exp(thp / (1.0 + xi[0]))) + kPi_2;
      *rho = (-1.0 - xi[0]) / cos(*th);
      break;
  }
}

#pragma clang fp contract(fast)

}  // namespace mfb
