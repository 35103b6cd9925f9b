/* Oracle: CPU restatement of the reference's 3D BEM assembly hot path (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library: it is
 * the checker for the HIP product and the timed CPU baseline.  The product never calls it.
 *
 * Every function below restates one routine of the reference (file:line given, paths relative to
 * the MultiFEBE tree).  Decisions (ball test, nearest point, quadrature-order estimators, subdivision
 * order) follow the reference step by step so that the same integration rule is chosen for every
 * pair.  The reference carries its nearest-point Newton iteration in real128; this restatement uses
 * long double (64-bit mantissa) for those internals.  The Gauss-Legendre tables are computed here
 * in long double and then rounded to 17 significant digits, which is the precision of the
 * reference's data statements (lib/fbem/src/resources_quad_rules/gl11.rc, gl01.rc).
 *
 * Parity pins: tests/test_oracle_golden.py checks every public function against vectors produced by
 * the compiled reference library (oracle/ref_harness.f90), and against closed forms (rigid-body
 * identity of the assembled H, Mantic C = I/2 at smooth points, E_n recurrences).
 */
#include "mfbem_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex zc;

/* numerical.f90:74-93 */
#define C_PI 3.14159265358979323846264338328
#define C_2PI 6.28318530717958623199592693709
#define C_PI_2 1.57079632679489661923132169164
#define C_PI_4 0.78539816339744830961566084582
#define C_1_4PI 0.07957747154594767280411105048
/* shape_functions.f90:280 */
#define XI_TOL 0.5e-12

static const double DKR[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

/* ------------------------------------------------------------------------------------------------
 * Gauss-Legendre tables (quad_rules.f90:79-136; resources_quad_rules/gl11.rc, gl01.rc)
 * ---------------------------------------------------------------------------------------------- */
static double GL11X[33][32], GL11W[33][32], GL01X[33][32], GL01W[33][32];
static int gl_ready = 0;

static double round17(long double v) {
  char buf[64];
  snprintf(buf, sizeof buf, "%.16Le", v);
  return strtod(buf, NULL);
}

static void gl_rule(int n, long double *x, long double *w) {
  /* Newton iteration on P_n in binary128, so that the 17-digit rounding below sees ~30 correct
   * digits. */
  typedef __float128 q_t;
  for (int i = 0; i < (n + 1) / 2; i++) {
    q_t z = (q_t)cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L));
    q_t p0, p1, dp = 1;
    for (int it = 0; it < 200; it++) {
      p0 = 1;
      p1 = z;
      for (int k = 1; k < n; k++) {
        q_t p2 = ((2 * k + 1) * z * p1 - k * p0) / (k + 1);
        p0 = p1;
        p1 = p2;
      }
      dp = (n == 1) ? (q_t)1 : n * (z * p1 - p0) / (z * z - 1);
      q_t dz = p1 / dp;
      z -= dz;
      if (dz < 0) dz = -dz;
      if (dz < (q_t)1e-32L) break;
    }
    p0 = 1;
    p1 = z;
    for (int k = 1; k < n; k++) {
      q_t p2 = ((2 * k + 1) * z * p1 - k * p0) / (k + 1);
      p0 = p1;
      p1 = p2;
    }
    dp = (n == 1) ? (q_t)1 : n * (z * p1 - p0) / (z * z - 1);
    q_t wi = 2 / ((1 - z * z) * dp * dp);
    x[i] = (long double)(-z);
    x[n - 1 - i] = (long double)z;
    w[i] = (long double)wi;
    w[n - 1 - i] = (long double)wi;
  }
  if (n % 2 == 1) x[n / 2] = 0;
}

static void gl_init(void) {
  if (gl_ready) return;
#pragma omp critical(orc_gl_init)
  {
    if (!gl_ready) {
      for (int n = 1; n <= 32; n++) {
        long double x[32], w[32];
        gl_rule(n, x, w);
        for (int i = 0; i < n; i++) {
          GL11X[n][i] = round17(x[i]);
          GL11W[n][i] = round17(w[i]);
          GL01X[n][i] = round17(0.5L * (1 + x[i]));
          GL01W[n][i] = round17(0.5L * w[i]);
        }
      }
      gl_ready = 1;
    }
  }
}

void orc_gl_tables(double *x11, double *w11, double *x01, double *w01) {
  gl_init();
  for (int n = 1; n <= 32; n++)
    for (int i = 0; i < 32; i++) {
      int o = (n - 1) * 32 + i;
      x11[o] = i < n ? GL11X[n][i] : 0;
      w11[o] = i < n ? GL11W[n][i] : 0;
      x01[o] = i < n ? GL01X[n][i] : 0;
      w01[o] = i < n ? GL01W[n][i] : 0;
    }
}

/* ------------------------------------------------------------------------------------------------
 * Shape functions (resources_shape_functions/phi_quad9.rc, dphidxi{1,2}_quad9.rc, phi_quad4.rc,
 * dphidxi{1,2}_quad4.rc, phi_line3.rc, dphidxi_line3.rc, phi_line2.rc), continuous case (delta=0).
 * ---------------------------------------------------------------------------------------------- */
static const double XI_Q9[9][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}, {0, -1},
                                   {1, 0},   {0, 1},  {-1, 0}, {0, 0}};
/* quad edges as line3 (shape_functions.f90:866-882): end, end, middle */
static const int EDGE_Q9[4][3] = {{0, 1, 4}, {1, 2, 5}, {2, 3, 6}, {3, 0, 7}};

static inline void phi_q9(const double *xi, double *p) {
  double a1 = xi[0], a2 = xi[1];
  double a3 = 0.25 * a1 * (a1 + 1.0), a4 = 0.25 * a1 * (a1 - 1.0);
  double a5 = a2 * (a2 + 1.0), a6 = a2 * (a2 - 1.0);
  double a7 = 1.0 - a1 * a1, a8 = 1.0 - a2 * a2;
  p[0] = a4 * a6; p[1] = a3 * a6; p[2] = a3 * a5; p[3] = a4 * a5;
  p[4] = 0.5 * a6 * a7; p[5] = 2.0 * a3 * a8; p[6] = 0.5 * a5 * a7; p[7] = 2.0 * a4 * a8;
  p[8] = a7 * a8;
}

static inline void dphi_q9(const double *xi, double *d1, double *d2) {
  {
    double a3 = 2.0 * xi[0] + 1.0, a4 = 2.0 * xi[0] - 1.0, a5 = xi[1] + 1.0, a6 = xi[1] - 1.0;
    double a7 = 0.25 * xi[1], a8 = a5 * a6, a9 = -0.5 * a8, a10 = -xi[0] * xi[1];
    d1[0] = a7 * a4 * a6; d1[1] = a7 * a3 * a6; d1[2] = a7 * a3 * a5; d1[3] = a7 * a4 * a5;
    d1[4] = a10 * a6; d1[5] = a9 * a3; d1[6] = a10 * a5; d1[7] = a9 * a4;
    d1[8] = 2.0 * xi[0] * a8;
  }
  {
    double a3 = 2.0 * xi[1] + 1.0, a4 = 2.0 * xi[1] - 1.0, a5 = xi[0] + 1.0, a6 = xi[0] - 1.0;
    double a7 = 0.25 * xi[0], a8 = a5 * a6, a9 = -0.5 * a8, a10 = -xi[0] * xi[1];
    d2[0] = a7 * a6 * a4; d2[1] = a7 * a5 * a4; d2[2] = a7 * a5 * a3; d2[3] = a7 * a6 * a3;
    d2[4] = a9 * a4; d2[5] = a10 * a5; d2[6] = a9 * a3; d2[7] = a10 * a6;
    d2[8] = 2.0 * xi[1] * a8;
  }
}

static inline void phi_q4(const double *xi, double *p) {
  double a3 = 0.25 * (1.0 + xi[0]), a4 = 0.25 * (1.0 - xi[0]);
  double a5 = 1.0 + xi[1], a6 = 1.0 - xi[1];
  p[0] = a4 * a6; p[1] = a3 * a6; p[2] = a3 * a5; p[3] = a4 * a5;
}

static inline void dphi_q4(const double *xi, double *d1, double *d2) {
  double b2 = 0.25 * xi[1], b3 = 0.25;
  d1[0] = b2 - b3; d1[1] = -d1[0]; d1[2] = b2 + b3; d1[3] = -d1[2];
  double c2 = 0.25 * xi[0];
  d2[0] = c2 - b3; d2[1] = -c2 - b3; d2[2] = -d2[1]; d2[3] = -d2[0];
}

static inline void phi_l3(double xi, double *p) {
  double a2 = 0.5 * xi, a3 = xi - 1.0, a4 = xi + 1.0;
  p[0] = a2 * a3; p[1] = a2 * a4; p[2] = -a3 * a4;
}

static inline void dphi_l3(double xi, double *d) {
  double a2 = xi, a3 = 0.5;
  d[0] = a2 - a3; d[1] = a2 + a3; d[2] = -2.0 * a2;
}

static inline void phi_l3_ld(long double xi, long double *p) {
  long double a2 = 0.5L * xi, a3 = xi - 1.0L, a4 = xi + 1.0L;
  p[0] = a2 * a3; p[1] = a2 * a4; p[2] = -a3 * a4;
}

static inline void dphi_l3_ld(long double xi, long double *d) {
  d[0] = xi - 0.5L; d[1] = xi + 0.5L; d[2] = -2.0L * xi;
}

static inline void phi_q9_ld(const double *xi, long double *p) {
  long double a1 = xi[0], a2 = xi[1];
  long double a3 = 0.25L * a1 * (a1 + 1), a4 = 0.25L * a1 * (a1 - 1);
  long double a5 = a2 * (a2 + 1), a6 = a2 * (a2 - 1);
  long double a7 = 1 - a1 * a1, a8 = 1 - a2 * a2;
  p[0] = a4 * a6; p[1] = a3 * a6; p[2] = a3 * a5; p[3] = a4 * a5;
  p[4] = 0.5L * a6 * a7; p[5] = 2 * a3 * a8; p[6] = 0.5L * a5 * a7; p[7] = 2 * a4 * a8;
  p[8] = a7 * a8;
}

static inline void dphi_q9_ld(const double *xi, long double *d1, long double *d2) {
  long double x1 = xi[0], x2 = xi[1];
  {
    long double a3 = 2 * x1 + 1, a4 = 2 * x1 - 1, a5 = x2 + 1, a6 = x2 - 1;
    long double a7 = 0.25L * x2, a8 = a5 * a6, a9 = -0.5L * a8, a10 = -x1 * x2;
    d1[0] = a7 * a4 * a6; d1[1] = a7 * a3 * a6; d1[2] = a7 * a3 * a5; d1[3] = a7 * a4 * a5;
    d1[4] = a10 * a6; d1[5] = a9 * a3; d1[6] = a10 * a5; d1[7] = a9 * a4; d1[8] = 2 * x1 * a8;
  }
  {
    long double a3 = 2 * x2 + 1, a4 = 2 * x2 - 1, a5 = x1 + 1, a6 = x1 - 1;
    long double a7 = 0.25L * x1, a8 = a5 * a6, a9 = -0.5L * a8, a10 = -x1 * x2;
    d2[0] = a7 * a6 * a4; d2[1] = a7 * a5 * a4; d2[2] = a7 * a5 * a3; d2[3] = a7 * a6 * a3;
    d2[4] = a9 * a4; d2[5] = a10 * a5; d2[6] = a9 * a3; d2[7] = a10 * a6; d2[8] = 2 * x2 * a8;
  }
}

static inline double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* x, T1, T2 of a quad9 at xi */
static inline void q9_geom(const double *X, const double *xi, double *x, double *T1, double *T2) {
  double p[9], d1[9], d2[9];
  phi_q9(xi, p);
  dphi_q9(xi, d1, d2);
  for (int c = 0; c < 3; c++) { x[c] = 0; T1[c] = 0; T2[c] = 0; }
  for (int k = 0; k < 9; k++)
    for (int c = 0; c < 3; c++) {
      x[c] += p[k] * X[3 * k + c];
      T1[c] += d1[k] * X[3 * k + c];
      T2[c] += d2[k] * X[3 * k + c];
    }
}

static inline void cross(const double *a, const double *b, double *n) {
  n[0] = a[1] * b[2] - a[2] * b[1];
  n[1] = a[2] * b[0] - a[0] * b[2];
  n[2] = a[0] * b[1] - a[1] * b[0];
}

/* ------------------------------------------------------------------------------------------------
 * Lengths and element data (geometry.f90:1469-1504 jacobian3d_1d, 2794-2882 length3d,
 * 3357-3390 characteristic_length, 3904-4235 element_ball; quasisingular_integration.f90:812-975)
 * ---------------------------------------------------------------------------------------------- */
static double jac_l3(const double *X, double xi) {
  double d[3], t[3] = {0, 0, 0};
  dphi_l3(xi, d);
  for (int k = 0; k < 3; k++)
    for (int c = 0; c < 3; c++) t[c] += d[k] * X[3 * k + c];
  return sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
}

static double length_l3(const double *X, double tol) {
  gl_init();
  double lt = tol <= 0 ? 1e-6 : tol;
  int step = 3;
  double length = jac_l3(X, GL11X[1][0]) * GL11W[1][0], length_old, err = lt + 1.0;
  int j = 1 + step;
  while (err >= lt) {
    length_old = length;
    length = 0;
    for (int i = 0; i < j; i++) length += jac_l3(X, GL11X[j][i]) * GL11W[j][i];
    err = fabs((length_old - length) / length);
    j += step;
    if (j >= 32) {
      length = 0;
      for (int i = 0; i < 31; i++) length += jac_l3(X, GL11X[31][i]) * GL11W[31][i];
      length_old = length;
      length = 0;
      for (int i = 0; i < 32; i++) length += jac_l3(X, GL11X[32][i]) * GL11W[32][i];
      break;
    }
  }
  return length;
}

static double clen_q9(const double *X, double tol) {
  double l = 0;
  for (int e = 0; e < 4; e++) {
    double xe[9];
    for (int k = 0; k < 3; k++)
      for (int c = 0; c < 3; c++) xe[3 * k + c] = X[3 * EDGE_Q9[e][k] + c];
    double le = length_l3(xe, tol);
    if (e == 0 || le > l) l = le;
  }
  return l;
}

static int phijac_ngp_q9(const double *X, double error_height) {
  gl_init();
  double leh = error_height <= 0 ? 1e-6 : (error_height < 1e-12 ? 1e-12 : error_height);
  double integral[9] = {0}, integral_old[9];
  int j = 1, cont = 1;
  while (cont) {
    memcpy(integral_old, integral, sizeof integral);
    memset(integral, 0, sizeof integral);
    for (int k1 = 0; k1 < j; k1++) {
      for (int k2 = 0; k2 < j; k2++) {
        double xi[2] = {GL11X[j][k1], GL11X[j][k2]}, d1[9], d2[9], p[9], T1[3] = {0}, T2[3] = {0},
               N[3];
        dphi_q9(xi, d1, d2);
        for (int l = 0; l < 9; l++)
          for (int c = 0; c < 3; c++) {
            T1[c] += d1[l] * X[3 * l + c];
            T2[c] += d2[l] * X[3 * l + c];
          }
        cross(T1, T2, N);
        double jac = sqrt(dot3(N, N));
        double jw = jac * GL11W[j][k1] * GL11W[j][k2];
        phi_q9(xi, p);
        for (int l = 0; l < 9; l++) integral[l] += (p[l] + 1.0) * jw;
      }
    }
    if (j == 1) {
      j++;
    } else {
      cont = 0;
      for (int l = 0; l < 9; l++) {
        if (fabs((integral_old[l] - integral[l]) / integral[l]) > leh) {
          cont = 1;
          j++;
          break;
        }
      }
      if (j == 33) cont = 0;
    }
  }
  return j - 1;
}

static void ball_q9(const double *X, int glp, double *centre, double *radius) {
  gl_init();
  double esize = 0, xm[3] = {0, 0, 0};
  for (int k1 = 0; k1 < glp; k1++)
    for (int k2 = 0; k2 < glp; k2++) {
      double xi[2] = {GL11X[glp][k1], GL11X[glp][k2]}, x[3], T1[3], T2[3], N[3];
      q9_geom(X, xi, x, T1, T2);
      cross(T1, T2, N);
      double jw = sqrt(dot3(N, N)) * GL11W[glp][k1] * GL11W[glp][k2];
      esize += jw;
      for (int c = 0; c < 3; c++) xm[c] += x[c] * jw;
    }
  for (int c = 0; c < 3; c++) centre[c] = xm[c] / esize;
  double r = 0;
  for (int k = 0; k < 9; k++) {
    double v[3] = {X[3 * k] - centre[0], X[3 * k + 1] - centre[1], X[3 * k + 2] - centre[2]};
    double t = sqrt(dot3(v, v));
    if (r < t) r = t;
  }
  for (int k1 = 0; k1 < glp; k1++)
    for (int k2 = 0; k2 < glp; k2++) {
      double xi[2] = {GL11X[glp][k1], GL11X[glp][k2]}, p[9], x[3] = {0, 0, 0};
      phi_q9(xi, p);
      for (int k = 0; k < 9; k++)
        for (int c = 0; c < 3; c++) x[c] += p[k] * X[3 * k + c];
      double v[3] = {x[0] - centre[0], x[1] - centre[1], x[2] - centre[2]};
      double t = sqrt(dot3(v, v));
      if (r < t) r = t;
    }
  *radius = r;
}

void orc_element_data(const double *x, double qsi_re, double *csize, int *n_phi, double *centre,
                      double *radius) {
  *csize = clen_q9(x, 1e-9);
  *n_phi = phijac_ngp_q9(x, qsi_re);
  ball_q9(x, *n_phi, centre, radius);
}

/* ------------------------------------------------------------------------------------------------
 * Quasi-singular rule estimators (quasisingular_integration.f90:181-400 tables and setup,
 * 402-556 standard, 558-715 Telles)
 * ---------------------------------------------------------------------------------------------- */
typedef struct {
  double dmin[3][8], dmax[3][8], a0[3][8], a1[3][8], b1[3][8];
  double tdmin[3][8], tdmax[3][8], ta0[3][8], ta1[3][8], tb1[3][8];
} qsp_t;

/* Fitted coefficients of the N(d) curves, [curve][coefficient]; the numbers are the reference's
 * published fit (quasisingular_integration.f90:181-292) and must be identical to reproduce its
 * decisions. */
static const double LNR_A0[3][3] = {{2.511371433253553e-1, -8.318131267946077e-2, -1.973643360180158e-3},
                                    {1.588915580092415e-1, -9.469173675969373e-2, -2.500979779162619e-3},
                                    {1.303338290205561e-1, -9.041069352366727e-2, -2.302123487272636e-3}};
static const double LNR_A1[3][3] = {{-8.301769338383449e-1, -7.458331284021328e-2, -1.992535165956408e-3},
                                    {-8.575272011921337e-1, -6.866562987114574e-2, -1.735469217262563e-3},
                                    {-5.785210418300009e-1, -3.107089548772586e-2, -4.204818399993008e-4}};
static const double LNR_B1[3][3] = {{-1.114783338380396e-1, -6.169581618278328e-2, -1.873698813959166e-3},
                                    {-3.918234478283774e-1, -6.478096580939043e-2, -1.657459305149584e-3},
                                    {-1.933844506192423e-1, -3.303326826061181e-2, -6.497990057394490e-4}};
static const double D1RN_A0[3][6] = {
    {2.443943976922257e-1, -7.845125775746743e-2, 3.895326800087784e-2, 8.783124822346978e-4, -1.719172638314362e-3, -1.575912106080523e-3},
    {1.855869272054077e-1, -8.441784317192081e-2, 3.992519918023225e-2, 1.003034542932314e-3, -2.012460287158036e-3, -1.596684136739970e-3},
    {1.964994120809251e-1, -7.533682025056084e-2, 6.300967117113785e-2, 1.453521059340246e-3, -1.653933461489553e-3, -2.613368441217468e-3}};
static const double D1RN_A1[3][6] = {
    {-8.421216556191472e-1, -8.778262996280104e-2, 1.679496012553309e-2, 6.522695414940627e-4, -2.649807213627071e-3, -8.472009602597869e-4},
    {-8.012899104725332e-1, -7.060561516342915e-2, 5.713299243166406e-3, -1.952251227204736e-4, -1.984455335909311e-3, -6.680823196695776e-4},
    {-5.891778953073477e-1, -4.515354064860566e-2, 5.728946548799128e-3, 1.106131193536092e-6, -1.179100966502172e-3, -5.355782068751523e-4}};
static const double D1RN_B1[3][6] = {
    {-2.693032219216048e-2, -6.138311391493901e-2, 2.273921394673689e-2, 9.789745450155518e-4, -2.030377872751224e-3, -9.059289455187508e-4},
    {-2.353515879579821e-1, -5.141856801187955e-2, 1.693100431819354e-2, -1.491381226735667e-5, -1.330369292483563e-3, -1.264704521827383e-3},
    {-9.828561879516765e-2, -3.074365737570111e-2, 1.306905121742232e-2, 6.065437465834727e-5, -7.853234287066991e-4, -1.025889829482388e-3}};
static const double T_LNR_A0[3][3] = {{2.5834148732821033e-01, -8.8865167808209675e-02, -2.3640194116454241e-03},
                                      {5.1283007739426911e-01, -1.1479138136155583e-02, 1.7167160355773211e-03},
                                      {4.1061632629507161e-01, -3.0669024691563598e-02, 5.1920314153891956e-04}};
static const double T_LNR_A1[3][3] = {{-9.1721308433764848e-01, -6.2837979237008845e-02, -1.1580362636310136e-03},
                                      {-1.0632414684959128e+00, -1.1633428696469847e-01, -4.4173578665990879e-03},
                                      {-1.1740312565911180e+00, -1.3946434974037192e-01, -5.4264705520737193e-03}};
static const double T_LNR_B1[3][3] = {{-4.9431525769444484e-01, -6.5628053690049357e-02, -1.2683068366630545e-03},
                                      {-9.1742055835524161e-01, -1.4448614058834844e-01, -5.5849846784841173e-03},
                                      {-1.1668939380105636e+00, -1.7686541251294283e-01, -6.9562836980345301e-03}};
static const double T_D1RN_A0[3][6] = {
    {2.9379042486857926e-01, -7.6518438657183094e-02, 3.0011332045426935e-02, 5.8900790667561719e-04, -1.7642732320416377e-03, -1.3634192833061319e-03},
    {1.1155167367380846e-01, -9.0043220465500617e-02, 3.9605615471752335e-02, 1.0669688965738889e-03, -2.2428131063082419e-03, -1.6030309471606837e-03},
    {9.5587534756324991e-02, -8.5841370957250140e-02, 5.9724590614136340e-02, 1.4101322853742803e-03, -2.0424217927592597e-03, -2.4264102214015559e-03}};
static const double T_D1RN_A1[3][6] = {
    {-9.8339088502232042e-01, -8.6761356304300027e-02, 2.6028760963988956e-02, 7.7677507397256337e-04, -2.3619290417702245e-03, -1.2891431252005461e-03},
    {-6.5856321275500784e-01, -4.7354561095619759e-02, 1.9423558756522371e-03, -2.9087437598165618e-04, -1.1859755938049318e-03, -4.5988734937654318e-04},
    {-2.4634605754204883e-01, 1.0051411705899942e-02, -1.1994978255365170e-02, -1.0022065157916291e-03, 7.9639034155470554e-04, -1.8103634920482468e-04}};
static const double T_D1RN_B1[3][6] = {
    {-4.5453529921363467e-01, -7.3941529478921839e-02, 3.7672240427426076e-02, 9.8276802658567980e-04, -1.8897770518175436e-03, -1.8079562352345093e-03},
    {-2.6610069546799336e-01, -3.6048387668689889e-02, 1.3288129149913631e-02, -1.8099231920766130e-04, -7.6015457149682889e-04, -1.2247286683263207e-03},
    {7.5605400079843565e-02, 2.4021121564322980e-02, -9.3693044556669332e-03, -1.3532246088221670e-03, 1.3564674989692248e-03, -7.4467072775070487e-04}};

static double poly1q(const double *c, double x) { return c[0] + c[1] * x + c[2] * (x * x); }
static double poly2q(const double *c, double x, double y) {
  return c[0] + c[1] * x + c[2] * y + c[3] * x * y + c[4] * (x * x) + c[5] * (y * y);
}

/* log10(30.) and log10(2.) are default-real (single precision) literals in the reference */
#define LOG10_30F ((double)(float)1.4771212547196624)
#define LOG10_2F ((double)(float)0.3010299956639812)

static void qsp_setup(double relative_error, qsp_t *p) {
  double l = log10(fabs(relative_error));
  if (l > -3.0) l = -3.0;
  if (l < -15.0) l = -15.0;
  for (int c = 0; c < 3; c++) {
    p->a0[c][0] = poly1q(LNR_A0[c], l);
    p->a1[c][0] = poly1q(LNR_A1[c], l);
    p->b1[c][0] = poly1q(LNR_B1[c], l);
    p->ta0[c][0] = poly1q(T_LNR_A0[c], l);
    p->ta1[c][0] = poly1q(T_LNR_A1[c], l);
    p->tb1[c][0] = poly1q(T_LNR_B1[c], l);
    for (int i = 1; i <= 7; i++) {
      p->a0[c][i] = poly2q(D1RN_A0[c], l, (double)i);
      p->a1[c][i] = poly2q(D1RN_A1[c], l, (double)i);
      p->b1[c][i] = poly2q(D1RN_B1[c], l, (double)i);
      p->ta0[c][i] = poly2q(T_D1RN_A0[c], l, (double)i);
      p->ta1[c][i] = poly2q(T_D1RN_A1[c], l, (double)i);
      p->tb1[c][i] = poly2q(T_D1RN_B1[c], l, (double)i);
    }
    for (int i = 0; i <= 7; i++) {
      p->dmin[c][i] = pow(10.0, (LOG10_30F - p->a0[c][i]) / (p->a1[c][i] - LOG10_30F * p->b1[c][i]));
      p->dmax[c][i] = pow(10.0, (LOG10_2F - p->a0[c][i]) / (p->a1[c][i] - LOG10_2F * p->b1[c][i]));
      p->tdmin[c][i] = pow(10.0, (LOG10_30F - p->ta0[c][i]) / (p->ta1[c][i] - LOG10_30F * p->tb1[c][i]));
      p->tdmax[c][i] = pow(10.0, (LOG10_2F - p->ta0[c][i]) / (p->ta1[c][i] - LOG10_2F * p->tb1[c][i]));
      if (p->tdmin[c][i] < 1e-6) p->tdmin[c][i] = 1e-6;
      if (p->tdmin[c][i] > p->tdmax[c][i]) p->tdmin[c][i] = 1e-6;
    }
  }
}

/* Shared body of fbem_qs_n_estimation_standard / _telles for line (ndim=1) and quad (ndim=2)
 * elements, without the optional orientation arguments (the SBIE calls do not pass them). */
static int qs_estimate(const qsp_t *p, int telles, int ndim, int f, double d, const double *barxi) {
  const double(*dmin)[8] = telles ? p->tdmin : p->dmin;
  const double(*dmax)[8] = telles ? p->tdmax : p->dmax;
  const double(*a0)[8] = telles ? p->ta0 : p->a0;
  const double(*a1)[8] = telles ? p->ta1 : p->a1;
  const double(*b1)[8] = telles ? p->tb1 : p->b1;
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
    /* d>2: the non-Telles N3 curve is used by both estimators */
    if (d >= p->dmax[2][f]) {
      n = 2;
    } else {
      double ld = log10(d);
      n = (int)lround(pow(10.0, (p->a0[2][f] + p->a1[2][f] * ld) / (1.0 + p->b1[2][f] * ld)));
    }
  }
  if (n > 30) n = 0;
  return n;
}

/* ------------------------------------------------------------------------------------------------
 * Telles transformation (telles_transformation.f90:77-232), fbem_f_any branch
 * ---------------------------------------------------------------------------------------------- */
static double telles_barr(double d) {
  double r = d < 3.0 ? d / (0.89039 * d + 0.32883) : 1.0;
  if (r > 1.0) r = 1.0;
  return r;
}

static void telles11(double bxi, double br, double *c) {
  double w = bxi / (1.0 + 2.0 * br);
  double p = 1.0 / (3.0 * (1.0 + 2.0 * br)) * (3.0 - 2.0 * br - 3.0 * w * bxi);
  double q = w / 2.0 * ((3.0 - 2.0 * br) / (1.0 + 2.0 * br) - 2.0 * (w * w) - 1.0);
  double R2 = sqrt(q * q + p * p * p);
  double R31 = -q + R2, R32 = -q - R2, d13 = 1.0 / 3.0;
  R31 = R31 >= 0 ? pow(R31, d13) : -pow(fabs(R31), d13);
  R32 = R32 >= 0 ? pow(R32, d13) : -pow(fabs(R32), d13);
  double g = R31 + R32 + w;
  double Q = 1.0 + 3.0 * (g * g);
  c[0] = (1.0 - br) / Q;
  c[1] = -3.0 * g * c[0];
  c[2] = (br + 3.0 * (g * g)) / Q;
  c[3] = -c[1];
}

static inline void telles_xj(const double *c, double g, double *xi, double *j) {
  *xi = c[0] * (g * g * g) + c[1] * (g * g) + c[2] * g + c[3];
  *j = 3.0 * c[0] * (g * g) + 2.0 * c[1] * g + c[2];
}

/* ------------------------------------------------------------------------------------------------
 * Nearest element point (geometry.f90:4263-4316 nodes, 4329-4440 and 4574-4787 sampling,
 * 5072-5150 Newton steps, 5160-5408 minimisation, 5519-5572 driver)
 * ---------------------------------------------------------------------------------------------- */
static int check_xi_q(const double *xi) {
  return !(xi[0] < -1.0 - XI_TOL || xi[0] > 1.0 + XI_TOL || xi[1] < -1.0 - XI_TOL ||
           xi[1] > 1.0 + XI_TOL);
}

static void nearest_nodes(int nnod, const double *X, const double *p, int *node, double *r) {
  double v[3] = {X[0] - p[0], X[1] - p[1], X[2] - p[2]};
  *r = sqrt(dot3(v, v));
  *node = 0;
  for (int k = 1; k < nnod; k++) {
    double w[3] = {X[3 * k] - p[0], X[3 * k + 1] - p[1], X[3 * k + 2] - p[2]};
    double t = sqrt(dot3(w, w));
    if (t < *r) { *r = t; *node = k; }
  }
}

static double dist_l3(const double *X, const double *p, double xi) {
  double ph[3], x[3] = {0, 0, 0};
  phi_l3(xi, ph);
  for (int k = 0; k < 3; k++)
    for (int c = 0; c < 3; c++) x[c] += ph[k] * X[3 * k + c];
  double v[3] = {x[0] - p[0], x[1] - p[1], x[2] - p[2]};
  return sqrt(dot3(v, v));
}

static void sample_l3(const double *X, const double *p, double *xi_out, double *r_out) {
  const int ns = 25, nr = 14, nit = 2;
  double xmin = 0, rmin = 0;
  for (int k = 0; k <= ns + 1; k++) {
    double xi = -1.0 + 2.0 / (double)(ns + 1) * (double)k;
    double r = dist_l3(X, p, xi);
    if (k == 0) { xmin = xi; rmin = r; }
    if (r < rmin) { xmin = xi; rmin = r; }
  }
  double xold = xmin, width = 2.0 / (double)(ns + 1);
  for (int l = 0; l < nit; l++) {
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

static double dist_q9(const double *X, const double *p, const double *xi) {
  double ph[9], x[3] = {0, 0, 0};
  phi_q9(xi, ph);
  for (int k = 0; k < 9; k++)
    for (int c = 0; c < 3; c++) x[c] += ph[k] * X[3 * k + c];
  double v[3] = {x[0] - p[0], x[1] - p[1], x[2] - p[2]};
  return sqrt(dot3(v, v));
}

static void sample_q9(const double *X, const double *p, double *xi_out, double *r_out) {
  const int ns = 25, nr = 14, nit = 2;
  double xm[2] = {0, 0}, rmin = 0;
  for (int k1 = 0; k1 <= ns + 1; k1++) {
    double xi[2];
    xi[0] = -1.0 + 2.0 * (double)k1 / (double)(ns + 1);
    for (int k2 = 0; k2 <= ns + 1; k2++) {
      xi[1] = -1.0 + 2.0 * (double)k2 / (double)(ns + 1);
      double r = dist_q9(X, p, xi);
      if (k1 == 0 && k2 == 0) { xm[0] = xi[0]; xm[1] = xi[1]; rmin = r; }
      if (r < rmin) { xm[0] = xi[0]; xm[1] = xi[1]; rmin = r; }
    }
  }
  double xo[2] = {xm[0], xm[1]}, width = 2.0 / (double)(ns + 1);
  for (int l = 0; l < nit; l++) {
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

static void newton_l3(const double *X, const double *xi_i, double b, double *nb) {
  long double ph[3], d[3], xb[3] = {0, 0, 0}, db[3] = {0, 0, 0};
  phi_l3_ld(b, ph);
  dphi_l3_ld(b, d);
  for (int k = 0; k < 3; k++)
    for (int c = 0; c < 3; c++) {
      xb[c] += ph[k] * X[3 * k + c];
      db[c] += d[k] * X[3 * k + c];
    }
  long double a[3], ab = 0, bb = 0;
  for (int c = 0; c < 3; c++) {
    a[c] = xb[c] - db[c] * b - xi_i[c];
    ab += a[c] * db[c];
    bb += db[c] * db[c];
  }
  *nb = (double)(-ab / bb);
}

static void nearest_l3(const double *X, double cl, const double *x_i, double *barxi, double *rmin,
                       double *d) {
  int node;
  nearest_nodes(3, X, x_i, &node, rmin);
  static const double XIN[3] = {-1.0, 1.0, 0.0};
  *barxi = XIN[node];
  *d = *rmin / cl;
  if (*d < 1.0) {
    double hist[22], err[21];
    hist[0] = (*barxi < -1.0 || *barxi > 1.0) ? 0.0 : *barxi;
    int k = 0, info = 0;
    while (info == 0) {
      newton_l3(X, x_i, hist[k], &hist[k + 1]);
      err[k] = fabs(hist[k + 1] - hist[k]) * 0.5;
      if (err[k] <= 1e-14) {
        *barxi = hist[k + 1];
        info = 1;
      } else if (k + 1 == 20) {
        *barxi = hist[k + 1];
        info = 2;
      } else if (k + 1 > 5 && err[k] > err[k - 2]) {
        *barxi = hist[k + 1];
        info = 3;
      } else {
        k++;
      }
    }
    if (info == 1) {
      if (*barxi < -1.0 || *barxi > 1.0) {
        *barxi = -1.0;
        double v0[3] = {X[0] - x_i[0], X[1] - x_i[1], X[2] - x_i[2]};
        double r0 = sqrt(dot3(v0, v0));
        double v1[3] = {X[3] - x_i[0], X[4] - x_i[1], X[5] - x_i[2]};
        if (sqrt(dot3(v1, v1)) < r0) *barxi = 1.0;
      }
      long double ph[3], r[3] = {0, 0, 0};
      phi_l3_ld(*barxi, ph);
      for (int kk = 0; kk < 3; kk++)
        for (int c = 0; c < 3; c++) r[c] += ph[kk] * X[3 * kk + c];
      for (int c = 0; c < 3; c++) r[c] -= x_i[c];
      *rmin = (double)sqrtl(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    } else {
      sample_l3(X, x_i, barxi, rmin);
    }
    *d = *rmin / cl;
  }
}

static void newton_q9(const double *X, const double *x_i, const double *b, double *nb) {
  long double ph[9], d1[9], d2[9], xb[3] = {0, 0, 0}, t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
  phi_q9_ld(b, ph);
  dphi_q9_ld(b, d1, d2);
  for (int k = 0; k < 9; k++)
    for (int c = 0; c < 3; c++) {
      xb[c] += ph[k] * X[3 * k + c];
      t1[c] += d1[k] * X[3 * k + c];
      t2[c] += d2[k] * X[3 * k + c];
    }
  long double a[3], bb = 0, cc = 0, bc = 0, ab = 0, ac = 0;
  for (int c = 0; c < 3; c++) a[c] = xb[c] - t1[c] * b[0] - t2[c] * b[1] - x_i[c];
  for (int c = 0; c < 3; c++) {
    bb += t1[c] * t1[c];
    cc += t2[c] * t2[c];
    bc += t1[c] * t2[c];
    ab += a[c] * t1[c];
    ac += a[c] * t2[c];
  }
  long double det = bb * cc - bc * bc;
  nb[0] = (double)(-(ab * cc - bc * ac) / det);
  nb[1] = (double)(-(bb * ac - ab * bc) / det);
}

/* returns info */
static int minimize_q9(const double *X, const double *x_i, double *barxi, double *rmin) {
  double hist[22][2], err[21];
  if (!check_xi_q(barxi)) {
    hist[0][0] = 0;
    hist[0][1] = 0;
  } else {
    hist[0][0] = barxi[0];
    hist[0][1] = barxi[1];
  }
  int k = 0, info = 0;
  while (info == 0) {
    newton_q9(X, x_i, hist[k], hist[k + 1]);
    double e1 = hist[k + 1][0] - hist[k][0], e2 = hist[k + 1][1] - hist[k][1];
    err[k] = sqrt(e1 * e1 + e2 * e2) * 0.5;
    if (err[k] <= 1e-14) {
      barxi[0] = hist[k + 1][0]; barxi[1] = hist[k + 1][1];
      info = 1;
    } else if (k + 1 == 20) {
      barxi[0] = hist[k + 1][0]; barxi[1] = hist[k + 1][1];
      info = 2;
    } else if (k + 1 > 5 && err[k] >= err[k - 2]) {
      barxi[0] = hist[k + 1][0]; barxi[1] = hist[k + 1][1];
      info = 3;
    } else {
      k++;
    }
  }
  if (info == 1) {
    if (!check_xi_q(barxi)) {
      double rme[4], bxe[4];
      for (int ke = 0; ke < 4; ke++) {
        double xe[9];
        for (int j = 0; j < 3; j++)
          for (int c = 0; c < 3; c++) xe[3 * j + c] = X[3 * EDGE_Q9[ke][j] + c];
        double cl = length_l3(xe, 1e-12), dd;
        nearest_l3(xe, cl, x_i, &bxe[ke], &rme[ke], &dd);
      }
      int kb = 0;
      for (int ke = 1; ke < 4; ke++)
        if (rme[ke] < rme[kb]) kb = ke;
      double ph[3];
      phi_l3(bxe[kb], ph);
      barxi[0] = 0;
      barxi[1] = 0;
      for (int j = 0; j < 3; j++) {
        barxi[0] += ph[j] * XI_Q9[EDGE_Q9[kb][j]][0];
        barxi[1] += ph[j] * XI_Q9[EDGE_Q9[kb][j]][1];
      }
    }
    long double ph[9], r[3] = {0, 0, 0};
    phi_q9_ld(barxi, ph);
    for (int kk = 0; kk < 9; kk++)
      for (int c = 0; c < 3; c++) r[c] += ph[kk] * X[3 * kk + c];
    for (int c = 0; c < 3; c++) r[c] -= x_i[c];
    *rmin = (double)sqrtl(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  }
  return info;
}

void orc_nearest_quad9(const double *X, double cl, const double *x_i, double *barxi, double *rmin,
                       double *d, int *method) {
  int node;
  nearest_nodes(9, X, x_i, &node, rmin);
  barxi[0] = XI_Q9[node][0];
  barxi[1] = XI_Q9[node][1];
  *d = *rmin / cl;
  if (*d < 1.0) {
    int info = minimize_q9(X, x_i, barxi, rmin);
    if (info != 1) {
      sample_q9(X, x_i, barxi, rmin);
      *method = 3;
    } else {
      *method = 2;
    }
    *d = *rmin / cl;
  } else {
    *method = 1;
  }
}

/* fbem_obtain_element_subdivision_coordinates (geometry.f90:2591-2697) for quad9 and line3 */
static void subdiv_q9(const double *X, const double xs[4][2], double *Xs) {
  for (int k = 0; k < 9; k++) {
    double ps[4], xi[2] = {0, 0}, ph[9];
    phi_q4(XI_Q9[k], ps);
    for (int i = 0; i < 4; i++) {
      xi[0] += ps[i] * xs[i][0];
      xi[1] += ps[i] * xs[i][1];
    }
    phi_q9(xi, ph);
    for (int c = 0; c < 3; c++) {
      Xs[3 * k + c] = 0;
      for (int i = 0; i < 9; i++) Xs[3 * k + c] += ph[i] * X[3 * i + c];
    }
  }
}

static void subdiv_l3(const double *X, const double *xs, double *Xs) {
  static const double XIN[3] = {-1.0, 1.0, 0.0};
  for (int k = 0; k < 3; k++) {
    double a1 = 0.5 * XIN[k];
    double ps0 = 0.5 - a1, ps1 = 0.5 + a1;
    double xi = 0;
    xi += ps0 * xs[0];
    xi += ps1 * xs[1];
    double ph[3];
    phi_l3(xi, ph);
    for (int c = 0; c < 3; c++) {
      Xs[3 * k + c] = 0;
      for (int i = 0; i < 3; i++) Xs[3 * k + c] += ph[i] * X[3 * i + c];
    }
  }
}

/* ------------------------------------------------------------------------------------------------
 * Kernels
 * ---------------------------------------------------------------------------------------------- */
typedef struct {
  int harmonic;
  double cteu1, cteu2, ctet1, ctet2; /* static (bem_staela3d.f90:616-619) */
  zc k1, k2, psi[6], chi[6], T1[10], T2[9], T3[9], cte_u, cte_t; /* bem_harela3d.f90:118-134 */
} kpar_t;

typedef struct {
  double hr[81], gr[81];
  zc hc[81], gc[81];
  long long npts;
} acc_t;

static void acc_zero(acc_t *a) {
  memset(a->hr, 0, sizeof a->hr);
  memset(a->gr, 0, sizeof a->gr);
  for (int i = 0; i < 81; i++) { a->hc[i] = 0; a->gc[i] = 0; }
  a->npts = 0;
}

static void acc_add(acc_t *a, const acc_t *b) {
  for (int i = 0; i < 81; i++) {
    a->hr[i] += b->hr[i]; a->gr[i] += b->gr[i];
    a->hc[i] += b->hc[i]; a->gc[i] += b->gc[i];
  }
  a->npts += b->npts;
}

static void kpar_static(double mu, double nu, kpar_t *k) {
  memset(k, 0, sizeof *k);
  k->cteu1 = 1.0 / (16.0 * C_PI * mu * (1.0 - nu));
  k->cteu2 = 3.0 - 4.0 * nu;
  k->ctet1 = -1.0 / (8.0 * C_PI * (1.0 - nu));
  k->ctet2 = 1.0 - 2.0 * nu;
}

/* fbem_bem_harela3d_calculate_parameters (bem_harela3d.f90:143-287), u* and t* parts */
static void kpar_harmonic(zc lambda, zc mu, double rho, double omega, kpar_t *k) {
  memset(k, 0, sizeof *k);
  k->harmonic = 1;
  const zc Iu = _Complex_I;
  zc c1 = csqrt((lambda + 2.0 * mu) / rho), c2 = csqrt(mu / rho);
  zc k1 = omega / c1, k2 = omega / c2;
  k->k1 = k1;
  k->k2 = k2;
  zc c2s = c2 * c2, c1s = c1 * c1, r = c2s / c1s;
  zc ik1 = Iu * k1, ik2 = Iu * k2;
  k->psi[0] = 0.5 * (1.0 + r);
  k->psi[1] = -1.0 / 3.0 * (2.0 / c2 + c2s / (c1s * c1)) * Iu * omega;
  k->psi[2] = Iu * k1 / (k2 * k2);
  k->psi[3] = 1.0 / (Iu * k2);
  k->psi[4] = 1.0 / (k2 * k2);
  k->psi[5] = -1.0 / (k2 * k2);
  k->chi[0] = -0.5 * (1.0 - r);
  k->chi[1] = -r;
  k->chi[2] = -3.0 * r / ik1;
  k->chi[3] = 3.0 / ik2;
  k->chi[4] = -3.0 * r / (ik1 * ik1);
  k->chi[5] = 3.0 / (ik2 * ik2);
  k->T1[0] = 3.0 * (r - 1.0);
  k->T1[1] = -0.25 * (1.0 / c2s - c2s / (c1s * c1s)) * (omega * omega);
  k->T1[2] = -2.0 * Iu * k1 * r;
  k->T1[3] = 2.0 * Iu * k2;
  k->T1[4] = -12.0 * r;
  k->T1[5] = 12.0;
  k->T1[6] = -r * 30.0 / ik1;
  k->T1[7] = 30.0 / ik2;
  k->T1[8] = -r * 30.0 / (ik1 * ik1);
  k->T1[9] = 30.0 / (ik2 * ik2);
  k->T2[0] = -r;
  k->T2[1] = -0.25 * (1.0 / c2s + c2s / (c1s * c1s)) * (omega * omega);
  k->T2[2] = -Iu * k2;
  k->T2[3] = 2.0 * r;
  k->T2[4] = -3.0;
  k->T2[5] = r * 6.0 / ik1;
  k->T2[6] = -6.0 / ik2;
  k->T2[7] = r * 6.0 / (ik1 * ik1);
  k->T2[8] = -6.0 / (ik2 * ik2);
  k->T3[0] = r;
  k->T3[1] = 0.25 * (3.0 * c2s / (c1s * c1s) - 2.0 / c1s + 1.0 / c2s) * (omega * omega);
  k->T3[2] = (2.0 * r - 1.0) * Iu * k1;
  k->T3[3] = 4.0 * r - 1.0;
  k->T3[4] = -2.0;
  k->T3[5] = r * 6.0 / ik1;
  k->T3[6] = -6.0 / ik2;
  k->T3[7] = r * 6.0 / (ik1 * ik1);
  k->T3[8] = -6.0 / (ik2 * ik2);
  k->cte_u = C_1_4PI / mu;
  k->cte_t = C_1_4PI;
}

/* fbem_zexp_decomposed (numerical.f90:1258-1330) for one argument */
static void zexp_dec(zc z, zc *E) {
  double az = cabs(z);
  if (az <= 1.0) {
    int n = az <= 1e-6 ? 1 : (int)lround(pow(10.0, 1.176 + 0.175 * log10(az)));
    zc term[22];
    term[0] = 1.0;
    term[1] = z;
    double fact = 1.0;
    zc zp = z;
    for (int k = 2; k <= n + 6; k++) {
      fact = fact * (double)k;
      zp = zp * z;
      term[k] = zp / fact;
    }
    E[6] = 0;
    for (int k = n + 6; k >= 6; k--) E[6] = E[6] + term[k];
    E[5] = E[6] + term[5];
    E[4] = E[5] + term[4];
    E[3] = E[4] + term[3];
    E[2] = E[3] + term[2];
    E[1] = E[2] + term[1];
    E[0] = E[1] + term[0];
  } else {
    zc ez = cexp(z), z2 = z * z, z3 = z2 * z, z4 = z3 * z, z5 = z4 * z;
    E[0] = ez;
    E[1] = ez - 1.0;
    E[2] = E[1] - z;
    E[3] = E[2] - 0.5 * z2;
    E[4] = E[3] - (1.0 / 6.0) * z3;
    E[5] = E[4] - (1.0 / 24.0) * z4;
    E[6] = E[5] - (1.0 / 120.0) * z5;
  }
}

void orc_zexp(double zr, double zi, double *E) {
  zc e[7];
  zexp_dec(zr + _Complex_I * zi, e);
  for (int k = 0; k < 7; k++) {
    E[2 * k] = creal(e[k]);
    E[2 * k + 1] = cimag(e[k]);
  }
}

/* The harmonic scalar factors psi, chi, TT1..TT3 at distance r (bem_harela3d.f90:660-682).
 * full=0 drops the 1/r^2 leading terms of TT1..TT3 (regular parts used by sbie_int). */
static void har_factors(const kpar_t *k, double r, int full, zc *psi, zc *chi, zc *TT1, zc *TT2,
                        zc *TT3) {
  const zc Iu = _Complex_I;
  double d1 = 1.0 / r, d2 = d1 * d1, d3 = d2 * d1, d4 = d3 * d1;
  zc z1 = -(Iu * k->k1 * r), z2 = -(Iu * k->k2 * r), E1[7], E2[7];
  zexp_dec(z1, E1);
  zexp_dec(z2, E2);
  E1[2] *= d1; E2[2] *= d1;
  E1[3] *= d2; E2[3] *= d2;
  E1[4] *= d3; E2[4] *= d3;
  E1[5] *= d4; E2[5] *= d4;
  *psi = k->psi[0] * d1 + k->psi[1] + E2[2] + k->psi[2] * E1[3] + k->psi[3] * E2[3] +
         k->psi[4] * E1[4] + k->psi[5] * E2[4];
  *chi = k->chi[0] * d1 + k->chi[1] * E1[2] + E2[2] + k->chi[2] * E1[3] + k->chi[3] * E2[3] +
         k->chi[4] * E1[4] + k->chi[5] * E2[4];
  zc t1 = k->T1[1] + k->T1[2] * E1[2] + k->T1[3] * E2[2] + k->T1[4] * E1[3] + k->T1[5] * E2[3] +
          k->T1[6] * E1[4] + k->T1[7] * E2[4] + k->T1[8] * E1[5] + k->T1[9] * E2[5];
  zc t2 = k->T2[1] + k->T2[2] * E2[2] + k->T2[3] * E1[3] + k->T2[4] * E2[3] + k->T2[5] * E1[4] +
          k->T2[6] * E2[4] + k->T2[7] * E1[5] + k->T2[8] * E2[5];
  zc t3 = k->T3[1] + k->T3[2] * E1[2] + k->T3[3] * E1[3] + k->T3[4] * E2[3] + k->T3[5] * E1[4] +
          k->T3[6] * E2[4] + k->T3[7] * E1[5] + k->T3[8] * E2[5];
  if (full) {
    *TT1 = k->T1[0] * d2 + t1;
    *TT2 = k->T2[0] * d2 + t2;
    *TT3 = k->T3[0] * d2 + t3;
  } else {
    *TT1 = t1;
    *TT2 = t2;
    *TT3 = t3;
  }
}

/* One exterior integration point (bem_staela3d.f90:636-645 / bem_harela3d.f90:657-690). */
static inline void ext_point(const kpar_t *k, const double *x, const double *n, const double *x_i,
                             const double *pjw, const double *sjw, acc_t *a) {
  double rv[3] = {x[0] - x_i[0], x[1] - x_i[1], x[2] - x_i[2]};
  double r = sqrt(dot3(rv, rv)), d1r = 1.0 / r, d1r2 = d1r * d1r;
  double drdx[3] = {rv[0] * d1r, rv[1] * d1r, rv[2] * d1r};
  double drdn = dot3(drdx, n);
  a->npts++;
  if (!k->harmonic) {
    for (int il = 0; il < 3; il++)
      for (int ik = 0; ik < 3; ik++) {
        double fs_u = d1r * (k->cteu2 * DKR[il][ik] + drdx[il] * drdx[ik]);
        double fs_t = d1r2 * ((k->ctet2 * DKR[il][ik] + 3.0 * drdx[il] * drdx[ik]) * drdn +
                              k->ctet2 * (n[il] * drdx[ik] - n[ik] * drdx[il]));
        double *h = a->hr + 9 * il + 27 * ik, *g = a->gr + 9 * il + 27 * ik;
        for (int m = 0; m < 9; m++) {
          h[m] += fs_t * pjw[m];
          g[m] += fs_u * sjw[m];
        }
      }
  } else {
    zc psi, chi, TT1, TT2, TT3;
    har_factors(k, r, 1, &psi, &chi, &TT1, &TT2, &TT3);
    for (int il = 0; il < 3; il++)
      for (int ik = 0; ik < 3; ik++) {
        zc fs_u = psi * DKR[il][ik] - chi * drdx[il] * drdx[ik];
        zc fs_t = TT1 * drdx[il] * drdx[ik] * drdn + TT2 * (drdn * DKR[il][ik] + drdx[ik] * n[il]) +
                  TT3 * drdx[il] * n[ik];
        zc *h = a->hc + 9 * il + 27 * ik, *g = a->gc + 9 * il + 27 * ik;
        for (int m = 0; m < 9; m++) {
          h[m] += fs_t * pjw[m];
          g[m] += fs_u * sjw[m];
        }
      }
  }
}

static void acc_finish(const kpar_t *k, int reverse, acc_t *a) {
  for (int i = 0; i < 81; i++) {
    if (!k->harmonic) {
      a->hr[i] *= k->ctet1;
      a->gr[i] *= k->cteu1;
      if (reverse) a->hr[i] = -a->hr[i];
    } else {
      a->hc[i] = k->cte_t * a->hc[i];
      a->gc[i] = k->cte_u * a->gc[i];
      if (reverse) a->hc[i] = -a->hc[i];
    }
  }
}

/* ------------------------------------------------------------------------------------------------
 * Element with precalculated datasets (bem_general.f90:51-107, 163-759 quad branch)
 * ---------------------------------------------------------------------------------------------- */
typedef struct {
  double X[27];
  double cl, centre[3], radius;
  int gln_far;
  int n_ps, ps_gln[30], ps_gln_max;
  double *ps_x[30], *ps_n[30], *ps_jw[30]; /* jw stored as 9 phi*j*w per point (p and s equal) */
} elem_t;

static void elem_init(elem_t *e, const double *X, double qsi_re, int n_ps, const int *ps_gln) {
  gl_init();
  memcpy(e->X, X, sizeof e->X);
  orc_element_data(X, qsi_re, &e->cl, &e->gln_far, e->centre, &e->radius);
  e->n_ps = n_ps;
  e->ps_gln_max = 0;
  for (int i = 0; i < n_ps; i++) {
    int g = ps_gln[i], ng = g * g;
    e->ps_gln[i] = g;
    if (g > e->ps_gln_max) e->ps_gln_max = g;
    e->ps_x[i] = malloc(sizeof(double) * 3 * ng);
    e->ps_n[i] = malloc(sizeof(double) * 3 * ng);
    e->ps_jw[i] = malloc(sizeof(double) * 9 * ng);
    for (int k1 = 0; k1 < g; k1++)
      for (int k2 = 0; k2 < g; k2++) {
        int kt = k2 + k1 * g;
        double xi[2] = {GL11X[g][k1], GL11X[g][k2]}, w1 = GL11W[g][k1], w2 = GL11W[g][k2];
        double x[3], T1[3], T2[3], N[3], p[9];
        q9_geom(X, xi, x, T1, T2);
        cross(T1, T2, N);
        double j = sqrt(dot3(N, N));
        phi_q9(xi, p);
        for (int c = 0; c < 3; c++) {
          e->ps_x[i][3 * kt + c] = x[c];
          e->ps_n[i][3 * kt + c] = N[c] / j;
        }
        for (int m = 0; m < 9; m++) e->ps_jw[i][9 * kt + m] = p[m] * j * w1 * w2;
      }
  }
}

static void elem_free(elem_t *e) {
  for (int i = 0; i < e->n_ps; i++) {
    free(e->ps_x[i]);
    free(e->ps_n[i]);
    free(e->ps_jw[i]);
  }
}

/* ext_pre (bem_staela3d.f90:592-650, bem_harela3d.f90:628-700) */
static void ext_pre(const kpar_t *k, const elem_t *e, int ps, const double *x_i, acc_t *a) {
  int ng = e->ps_gln[ps] * e->ps_gln[ps];
  for (int kip = 0; kip < ng; kip++)
    ext_point(k, e->ps_x[ps] + 3 * kip, e->ps_n[ps] + 3 * kip, x_i, e->ps_jw[ps] + 9 * kip,
              e->ps_jw[ps] + 9 * kip, a);
}

/* ext_st, quadrilateral branch (bem_staela3d.f90:654-969, bem_harela3d.f90:702-1048) */
static void ext_st(const kpar_t *k, const elem_t *e, const double xs[4][2], const double *x_i,
                   const double *barxip, double barr, int gln, acc_t *a) {
  double tp1[4], tp2[4];
  telles11(barxip[0], barr, tp1);
  telles11(barxip[1], barr, tp2);
  for (int k1 = 0; k1 < gln; k1++) {
    double g1 = GL11X[gln][k1], w1 = GL11W[gln][k1], xp[2], jt1, jt2;
    telles_xj(tp1, g1, &xp[0], &jt1);
    for (int k2 = 0; k2 < gln; k2++) {
      double g2 = GL11X[gln][k2], w2 = GL11W[gln][k2];
      telles_xj(tp2, g2, &xp[1], &jt2);
      double gp[4], d1[4], d2[4];
      phi_q4(xp, gp);
      dphi_q4(xp, d1, d2);
      double xi[2] = {0, 0}, dx1[2] = {0, 0}, dx2[2] = {0, 0};
      for (int kk = 0; kk < 4; kk++) {
        xi[0] += gp[kk] * xs[kk][0]; xi[1] += gp[kk] * xs[kk][1];
        dx1[0] += d1[kk] * xs[kk][0]; dx1[1] += d1[kk] * xs[kk][1];
        dx2[0] += d2[kk] * xs[kk][0]; dx2[1] += d2[kk] * xs[kk][1];
      }
      double js = dx1[0] * dx2[1] - dx1[1] * dx2[0];
      double x[3], T1[3], T2[3], N[3], n[3], p[9], pjw[9];
      q9_geom(e->X, xi, x, T1, T2);
      cross(T1, T2, N);
      double jg = sqrt(dot3(N, N));
      for (int c = 0; c < 3; c++) n[c] = N[c] / jg;
      double jw = jg * js * jt1 * jt2 * w1 * w2;
      phi_q9(xi, p);
      for (int m = 0; m < 9; m++) pjw[m] = p[m] * jw;
      ext_point(k, x, n, x_i, pjw, pjw, a);
    }
  }
}

/* ext_adp (bem_staela3d.f90:973-1096, bem_harela3d.f90:1050-1172), recursive 4-way subdivision */
static void ext_adp(const kpar_t *k, const elem_t *e, double xs[4][2], const double *x_i,
                    const qsp_t *qsp, int ks, int ns, acc_t *a, long long *leaves) {
  double barxip[2], rmin, d;
  int method;
  if (ks == 1) {
    xs[0][0] = -1; xs[0][1] = -1; xs[1][0] = 1; xs[1][1] = -1;
    xs[2][0] = 1; xs[2][1] = 1; xs[3][0] = -1; xs[3][1] = 1;
    orc_nearest_quad9(e->X, e->cl, x_i, barxip, &rmin, &d, &method);
  } else {
    double Xs[27];
    subdiv_q9(e->X, (const double(*)[2])xs, Xs);
    double cl = clen_q9(Xs, 1e-12);
    orc_nearest_quad9(Xs, cl, x_i, barxip, &rmin, &d, &method);
  }
  int gln_near = qs_estimate(qsp, 1, 2, 5, d, barxip);
  int subdivide = 0;
  if (ks == ns) {
    if (gln_near == 0) gln_near = 30;
  } else if (gln_near == 0) {
    subdivide = 1;
  }
  if (subdivide) {
    double t[4][2], c0 = 0.25 * (xs[0][0] + xs[1][0] + xs[2][0] + xs[3][0]),
                    c1 = 0.25 * (xs[0][1] + xs[1][1] + xs[2][1] + xs[3][1]);
    /* subquad 1 */
    t[0][0] = xs[0][0]; t[0][1] = xs[0][1];
    t[1][0] = 0.5 * (xs[0][0] + xs[1][0]); t[1][1] = 0.5 * (xs[0][1] + xs[1][1]);
    t[2][0] = c0; t[2][1] = c1;
    t[3][0] = 0.5 * (xs[0][0] + xs[3][0]); t[3][1] = 0.5 * (xs[0][1] + xs[3][1]);
    ext_adp(k, e, t, x_i, qsp, ks + 1, ns, a, leaves);
    /* subquad 2 */
    t[0][0] = 0.5 * (xs[0][0] + xs[1][0]); t[0][1] = 0.5 * (xs[0][1] + xs[1][1]);
    t[1][0] = xs[1][0]; t[1][1] = xs[1][1];
    t[2][0] = 0.5 * (xs[1][0] + xs[2][0]); t[2][1] = 0.5 * (xs[1][1] + xs[2][1]);
    t[3][0] = c0; t[3][1] = c1;
    ext_adp(k, e, t, x_i, qsp, ks + 1, ns, a, leaves);
    /* subquad 3 */
    t[0][0] = c0; t[0][1] = c1;
    t[1][0] = 0.5 * (xs[1][0] + xs[2][0]); t[1][1] = 0.5 * (xs[1][1] + xs[2][1]);
    t[2][0] = xs[2][0]; t[2][1] = xs[2][1];
    t[3][0] = 0.5 * (xs[2][0] + xs[3][0]); t[3][1] = 0.5 * (xs[2][1] + xs[3][1]);
    ext_adp(k, e, t, x_i, qsp, ks + 1, ns, a, leaves);
    /* subquad 4 */
    t[0][0] = 0.5 * (xs[0][0] + xs[3][0]); t[0][1] = 0.5 * (xs[0][1] + xs[3][1]);
    t[1][0] = c0; t[1][1] = c1;
    t[2][0] = 0.5 * (xs[2][0] + xs[3][0]); t[2][1] = 0.5 * (xs[2][1] + xs[3][1]);
    t[3][0] = xs[3][0]; t[3][1] = xs[3][1];
    ext_adp(k, e, t, x_i, qsp, ks + 1, ns, a, leaves);
  } else {
    double barr = telles_barr(d);
    int gln = gln_near > e->gln_far ? gln_near : e->gln_far;
    acc_t tmp;
    acc_zero(&tmp);
    ext_st(k, e, (const double(*)[2])xs, x_i, barxip, barr, gln, &tmp);
    acc_add(a, &tmp);
    if (leaves) (*leaves)++;
  }
}

/* ------------------------------------------------------------------------------------------------
 * Singular integration (bem_staela3d.f90:1100-1381, bem_harela3d.f90:1174-1472) with the polar
 * transformation (polar_transformation.f90:251-546) and Stokes line integrals (:2245-2376)
 * ---------------------------------------------------------------------------------------------- */
static int check_xi_edge_q(const double *xi) {
  if (xi[0] < -1.0 - XI_TOL || xi[0] > 1.0 + XI_TOL || xi[1] < -1.0 - XI_TOL ||
      xi[1] > 1.0 + XI_TOL)
    return 0;
  return (xi[0] < -1.0 + XI_TOL || xi[0] > 1.0 - XI_TOL || xi[1] < -1.0 + XI_TOL ||
          xi[1] > 1.0 - XI_TOL);
}

static void polar_setup_q(const double *xi, int *nst, int *st, double th[8][2], double thp[8][2]) {
  int n = 0;
  if (!check_xi_edge_q(xi)) {
    n = 8;
    for (int i = 0; i < 8; i++) st[i] = i + 1;
  } else {
    const double t = XI_TOL;
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
  for (int k = 0; k < n; k++) {
    double a, b, th1, th2, th3, th4;
    switch (st[k]) {
      case 1:
        th1 = C_PI - asin((-1.0 - xi[1]) / sqrt((-1.0 - xi[0]) * (-1.0 - xi[0]) + (-1.0 - xi[1]) * (-1.0 - xi[1])));
        a = th1; b = 1.5 * C_PI;
        thp[k][0] = (1.0 + xi[1]) * log(tan(0.5 * (a - C_PI)));
        thp[k][1] = (1.0 + xi[1]) * log(tan(0.5 * (b - C_PI)));
        break;
      case 2:
        th2 = C_2PI + asin((-1.0 - xi[1]) / sqrt((1.0 - xi[0]) * (1.0 - xi[0]) + (-1.0 - xi[1]) * (-1.0 - xi[1])));
        a = 1.5 * C_PI; b = th2;
        thp[k][0] = (1.0 + xi[1]) * log(tan(0.5 * (a - C_PI)));
        thp[k][1] = (1.0 + xi[1]) * log(tan(0.5 * (b - C_PI)));
        break;
      case 3:
        th2 = C_2PI + asin((-1.0 - xi[1]) / sqrt((1.0 - xi[0]) * (1.0 - xi[0]) + (-1.0 - xi[1]) * (-1.0 - xi[1])));
        a = th2; b = C_2PI;
        thp[k][0] = (1.0 - xi[0]) * log(tan(0.5 * (a + C_PI_2)));
        thp[k][1] = (1.0 - xi[0]) * log(tan(0.5 * (b + C_PI_2)));
        break;
      case 4:
        th3 = asin((1.0 - xi[1]) / sqrt((1.0 - xi[0]) * (1.0 - xi[0]) + (1.0 - xi[1]) * (1.0 - xi[1])));
        a = 0.0; b = th3;
        thp[k][0] = (1.0 - xi[0]) * log(tan(0.5 * (a + C_PI_2)));
        thp[k][1] = (1.0 - xi[0]) * log(tan(0.5 * (b + C_PI_2)));
        break;
      case 5:
        th3 = asin((1.0 - xi[1]) / sqrt((1.0 - xi[0]) * (1.0 - xi[0]) + (1.0 - xi[1]) * (1.0 - xi[1])));
        a = th3; b = C_PI_2;
        thp[k][0] = (1.0 - xi[1]) * log(tan(0.5 * a));
        thp[k][1] = (1.0 - xi[1]) * log(tan(0.5 * b));
        break;
      case 6:
        th4 = C_PI - asin((1.0 - xi[1]) / sqrt((-1.0 - xi[0]) * (-1.0 - xi[0]) + (1.0 - xi[1]) * (1.0 - xi[1])));
        a = C_PI_2; b = th4;
        thp[k][0] = (1.0 - xi[1]) * log(tan(0.5 * a));
        thp[k][1] = (1.0 - xi[1]) * log(tan(0.5 * b));
        break;
      case 7:
        th4 = C_PI - asin((1.0 - xi[1]) / sqrt((-1.0 - xi[0]) * (-1.0 - xi[0]) + (1.0 - xi[1]) * (1.0 - xi[1])));
        a = th4; b = C_PI;
        thp[k][0] = (1.0 + xi[0]) * log(tan(0.5 * (a - C_PI_2)));
        thp[k][1] = (1.0 + xi[0]) * log(tan(0.5 * (b - C_PI_2)));
        break;
      default: /* 8 */
        th1 = C_PI - asin((-1.0 - xi[1]) / sqrt((-1.0 - xi[0]) * (-1.0 - xi[0]) + (-1.0 - xi[1]) * (-1.0 - xi[1])));
        a = C_PI; b = th1;
        thp[k][0] = (1.0 + xi[0]) * log(tan(0.5 * (a - C_PI_2)));
        thp[k][1] = (1.0 + xi[0]) * log(tan(0.5 * (b - C_PI_2)));
        break;
    }
    th[k][0] = a;
    th[k][1] = b;
  }
  *nst = n;
}

static void polar_angular_q(const double *xi, int st, double thp, double *th, double *rho) {
  switch (st) {
    case 1: case 2:
      *th = 2.0 * atan(exp(thp / (1.0 + xi[1]))) + C_PI;
      *rho = (-1.0 - xi[1]) / sin(*th);
      break;
    case 3: case 4:
      *th = 2.0 * atan(exp(thp / (1.0 - xi[0]))) - C_PI_2;
      *rho = (1.0 - xi[0]) / cos(*th);
      break;
    case 5: case 6:
      *th = 2.0 * atan(exp(thp / (1.0 - xi[1])));
      *rho = (1.0 - xi[1]) / sin(*th);
      break;
    default:
      *th = 2.0 * atan(exp(thp / (1.0 + xi[0]))) + C_PI_2;
      *rho = (-1.0 - xi[0]) / cos(*th);
      break;
  }
}

/* fbem_bem_staela3d_sbie_int_li: line integral eps_ijk e_k . t / r over a line3 edge */
static void int_li(const double *Xl, double *xs, const double *x_i, int ngp_min, const qsp_t *qsp,
                   int ks, int ns, double *hli) {
  double Xs[9];
  if (ks == 1) {
    xs[0] = -1.0;
    xs[1] = 1.0;
    memcpy(Xs, Xl, sizeof Xs);
  } else {
    subdiv_l3(Xl, xs, Xs);
  }
  double cl = length_l3(Xs, 1e-12), barxip, rmin, d;
  nearest_l3(Xs, cl, x_i, &barxip, &rmin, &d);
  int gln_near = qs_estimate(qsp, 1, 1, 1, d, &barxip);
  int subdivide = 0;
  if (ks == ns) {
    if (gln_near == 0) gln_near = 30;
  } else if (gln_near == 0) {
    subdivide = 1;
  }
  if (subdivide) {
    double t[2];
    t[0] = xs[0];
    t[1] = 0.5 * (xs[0] + xs[1]);
    int_li(Xl, t, x_i, ngp_min, qsp, ks + 1, ns, hli);
    t[0] = 0.5 * (xs[0] + xs[1]);
    t[1] = xs[1];
    int_li(Xl, t, x_i, ngp_min, qsp, ks + 1, ns, hli);
  } else {
    double tmp[9] = {0};
    int gln = gln_near > ngp_min ? gln_near : ngp_min;
    double barr = telles_barr(d), tp[4];
    telles11(barxip, barr, tp);
    for (int kip = 0; kip < gln; kip++) {
      double g = GL11X[gln][kip], w = GL11W[gln][kip], xip, jt;
      telles_xj(tp, g, &xip, &jt);
      double xi = 0.5 * (1.0 - xip) * xs[0] + 0.5 * (1.0 + xip) * xs[1];
      double js = 0.5 * (xs[1] - xs[0]);
      double ph[3], dp[3], x[3] = {0, 0, 0}, T[3] = {0, 0, 0};
      phi_l3(xi, ph);
      dphi_l3(xi, dp);
      for (int kk = 0; kk < 3; kk++)
        for (int c = 0; c < 3; c++) {
          x[c] += ph[kk] * Xl[3 * kk + c];
          T[c] += dp[kk] * Xl[3 * kk + c];
        }
      double jg = sqrt(dot3(T, T)), t[3] = {T[0] / jg, T[1] / jg, T[2] / jg};
      double rv[3] = {x[0] - x_i[0], x[1] - x_i[1], x[2] - x_i[2]};
      double r = sqrt(dot3(rv, rv)), dr1 = 1.0 / r;
      double jw = jg * js * jt * w;
      tmp[0 + 3 * 1] = tmp[0 + 3 * 1] - dr1 * t[2] * jw; /* (1,2) */
      tmp[0 + 3 * 2] = tmp[0 + 3 * 2] + dr1 * t[1] * jw; /* (1,3) */
      tmp[1 + 3 * 2] = tmp[1 + 3 * 2] - dr1 * t[0] * jw; /* (2,3) */
    }
    tmp[1 + 3 * 0] = -tmp[0 + 3 * 1];
    tmp[2 + 3 * 0] = -tmp[0 + 3 * 2];
    tmp[2 + 3 * 1] = -tmp[1 + 3 * 2];
    for (int i = 0; i < 9; i++) hli[i] += tmp[i];
  }
}

static void sbie_int(const kpar_t *k, const double *X, int reverse, const double *xi_i, acc_t *a) {
  gl_init();
  double pg[9], x_i[3] = {0, 0, 0}, pf_i[9];
  phi_q9(xi_i, pg);
  for (int m = 0; m < 9; m++)
    for (int c = 0; c < 3; c++) x_i[c] += pg[m] * X[3 * m + c];
  phi_q9(xi_i, pf_i);
  int nst, st[8];
  double th[8][2], thp[8][2];
  polar_setup_q(xi_i, &nst, st, th, thp);
  for (int ks = 0; ks < nst; ks++) {
    double thi = th[ks][0], thf = th[ks][1], thpi = thp[ks][0], thpf = thp[ks][1];
    int ngr = 15, ngt = 5 + (int)lround(25.0 * (thf - thi) / C_PI_2);
    for (int kt = 0; kt < ngt; kt++) {
      double tpp = GL01X[ngt][kt], wa = GL01W[ngt][kt], jtp = thpf - thpi;
      double tp = jtp * tpp + thpi, theta, rhoij;
      polar_angular_q(xi_i, st[ks], tp, &theta, &rhoij);
      double ct = cos(theta), sn = sin(theta);
      for (int kr = 0; kr < ngr; kr++) {
        double rp = GL01X[ngr][kr], wr = GL01W[ngr][kr], rho = rhoij * rp;
        double xi[2] = {xi_i[0] + rho * ct, xi_i[1] + rho * sn};
        double x[3], T1[3], T2[3], N[3], n[3], p[9];
        q9_geom(X, xi, x, T1, T2);
        cross(T1, T2, N);
        double jg = sqrt(dot3(N, N));
        for (int c = 0; c < 3; c++) n[c] = N[c] / jg;
        double rv[3] = {x[0] - x_i[0], x[1] - x_i[1], x[2] - x_i[2]};
        double r = sqrt(dot3(rv, rv)), d1r = 1.0 / r, d1r2 = d1r * d1r;
        double drdx[3] = {rv[0] * d1r, rv[1] * d1r, rv[2] * d1r}, drdn = dot3(drdx, n);
        double jw = jg * rho * jtp * wa * wr;
        phi_q9(xi, p);
        double pjw[9], dpjw[9];
        for (int m = 0; m < 9; m++) {
          pjw[m] = p[m] * jw;
          dpjw[m] = (p[m] - pf_i[m]);
        }
        a->npts++;
        if (!k->harmonic) {
          for (int il = 0; il < 3; il++)
            for (int ik = 0; ik < 3; ik++) {
              double fs_u = d1r * (k->cteu2 * DKR[il][ik] + drdx[il] * drdx[ik]);
              double fs_t = d1r2 * drdn * (k->ctet2 * DKR[il][ik] + 3.0 * drdx[il] * drdx[ik]);
              double fs_c = d1r2 * k->ctet2 * (n[il] * drdx[ik] - n[ik] * drdx[il]);
              double *h = a->hr + 9 * il + 27 * ik, *g = a->gr + 9 * il + 27 * ik;
              for (int m = 0; m < 9; m++) {
                h[m] += fs_t * pjw[m];
                g[m] += fs_u * pjw[m];
                h[m] += fs_c * dpjw[m] * jw;
              }
            }
        } else {
          zc psi, chi, TT1, TT2, TT3;
          har_factors(k, r, 0, &psi, &chi, &TT1, &TT2, &TT3);
          for (int il = 0; il < 3; il++)
            for (int ik = 0; ik < 3; ik++) {
              zc fs_u = psi * DKR[il][ik] - chi * drdx[il] * drdx[ik];
              zc fs_t = (TT1 + k->T1[0] * d1r2) * drdx[il] * drdx[ik] * drdn +
                        (TT2 + k->T2[0] * d1r2) * drdn * DKR[il][ik] + TT2 * drdx[ik] * n[il] +
                        TT3 * drdx[il] * n[ik];
              zc fs_c = k->T2[0] * d1r2 * (n[il] * drdx[ik] - n[ik] * drdx[il]);
              zc *h = a->hc + 9 * il + 27 * ik, *g = a->gc + 9 * il + 27 * ik;
              for (int m = 0; m < 9; m++) {
                h[m] += fs_t * pjw[m];
                g[m] += fs_u * pjw[m];
                h[m] += fs_c * dpjw[m] * jw;
              }
            }
        }
      }
    }
  }
  /* line integrals over the edges that bound an integrated sub-triangle */
  double hli[9] = {0};
  qsp_t q;
  qsp_setup(1e-15, &q);
  for (int ke = 0; ke < 4; ke++) {
    int integrate = 0;
    for (int s = 0; s < nst; s++)
      if (st[s] == 2 * (ke + 1) - 1 || st[s] == 2 * (ke + 1)) { integrate = 1; break; }
    if (integrate) {
      double Xl[9], xs[2];
      for (int j = 0; j < 3; j++)
        for (int c = 0; c < 3; c++) Xl[3 * j + c] = X[3 * EDGE_Q9[ke][j] + c];
      int_li(Xl, xs, x_i, 5, &q, 1, 16, hli);
    }
  }
  for (int il = 0; il < 3; il++)
    for (int ik = 0; ik < 3; ik++)
      for (int m = 0; m < 9; m++) {
        if (!k->harmonic)
          a->hr[m + 9 * il + 27 * ik] += pf_i[m] * k->ctet2 * hli[il + 3 * ik];
        else
          a->hc[m + 9 * il + 27 * ik] += pf_i[m] * k->T2[0] * hli[il + 3 * ik];
      }
  acc_finish(k, reverse, a);
}

/* ------------------------------------------------------------------------------------------------
 * Automatic integration of one pair (bem_staela3d.f90:524-589, bem_harela3d.f90:1474-1538)
 * ---------------------------------------------------------------------------------------------- */
typedef struct {
  int delta, gln_near, gln, ps;
  double barxi[2], d;
  long long leaves;
} trace_t;

static void sbie_auto(const kpar_t *k, const elem_t *e, int reverse, const double *x_i,
                      const qsp_t *qsp, int ns, acc_t *a, trace_t *tr) {
  double r[3] = {e->centre[0] - x_i[0], e->centre[1] - x_i[1], e->centre[2] - x_i[2]};
  double rmin = sqrt(dot3(r, r)) - e->radius, barxi[2], d;
  int delta;
  acc_zero(a);
  tr->gln_near = -1;
  tr->gln = -1;
  tr->ps = -1;
  tr->leaves = 0;
  if (rmin > 4.0 * e->radius) {
    delta = 0;
    barxi[0] = barxi[1] = 0;
    d = rmin / e->cl;
  } else {
    int method;
    orc_nearest_quad9(e->X, e->cl, x_i, barxi, &rmin, &d, &method);
    delta = d <= 1e-12 ? 1 : 0;
  }
  tr->delta = delta;
  tr->barxi[0] = barxi[0];
  tr->barxi[1] = barxi[1];
  tr->d = d;
  if (delta == 1) {
    sbie_int(k, e->X, reverse, barxi, a);
    return;
  }
  int gln_near = qs_estimate(qsp, 0, 2, 5, d, barxi);
  int gln = e->gln_far > gln_near ? e->gln_far : gln_near;
  tr->gln_near = gln_near;
  tr->gln = gln;
  if (gln <= e->ps_gln_max && gln_near > 0) {
    int ps = 0;
    for (int i = 0; i < e->n_ps; i++)
      if (e->ps_gln[i] >= gln) { ps = i; break; }
    tr->ps = ps + 1;
    ext_pre(k, e, ps, x_i, a);
    acc_finish(k, reverse, a);
  } else {
    tr->ps = 0;
    double xs[4][2];
    ext_adp(k, e, xs, x_i, qsp, 1, ns, a, &tr->leaves);
    acc_finish(k, reverse, a);
  }
}

static void pair_common(const kpar_t *k, const double *x, const double *x_i, double qsi_re, int ns,
                        int reverse, int n_ps, const int *ps_gln, int *trace, double *dtrace,
                        acc_t *a) {
  elem_t e;
  qsp_t q;
  trace_t tr;
  qsp_setup(qsi_re, &q);
  elem_init(&e, x, qsi_re, n_ps, ps_gln);
  sbie_auto(k, &e, reverse, x_i, &q, ns, a, &tr);
  elem_free(&e);
  if (trace) {
    trace[0] = tr.delta;
    trace[1] = tr.gln_near;
    trace[2] = tr.gln;
    trace[3] = tr.ps;
  }
  if (dtrace) {
    dtrace[0] = tr.barxi[0];
    dtrace[1] = tr.barxi[1];
    dtrace[2] = tr.d;
    dtrace[3] = (double)tr.leaves;
  }
}

int orc_pair_sta(const double *x, const double *x_i, double mu, double nu, double qsi_re, int ns,
                 int reverse, int n_ps, const int *ps_gln, int *trace, double *dtrace, double *h,
                 double *g) {
  kpar_t k;
  acc_t a;
  kpar_static(mu, nu, &k);
  pair_common(&k, x, x_i, qsi_re, ns, reverse, n_ps, ps_gln, trace, dtrace, &a);
  memcpy(h, a.hr, sizeof a.hr);
  memcpy(g, a.gr, sizeof a.gr);
  return 0;
}

int orc_pair_har(const double *x, const double *x_i, double lam_r, double lam_i, double mu_r,
                 double mu_i, double rho, double omega, double qsi_re, int ns, int reverse, int n_ps,
                 const int *ps_gln, int *trace, double *dtrace, double *h, double *g) {
  kpar_t k;
  acc_t a;
  kpar_harmonic(lam_r + _Complex_I * lam_i, mu_r + _Complex_I * mu_i, rho, omega, &k);
  pair_common(&k, x, x_i, qsi_re, ns, reverse, n_ps, ps_gln, trace, dtrace, &a);
  for (int i = 0; i < 81; i++) {
    h[2 * i] = creal(a.hc[i]);
    h[2 * i + 1] = cimag(a.hc[i]);
    g[2 * i] = creal(a.gc[i]);
    g[2 * i + 1] = cimag(a.gc[i]);
  }
  return 0;
}

/* ------------------------------------------------------------------------------------------------
 * Fixed rule (g, s): Gauss-Legendre order g on each of the 2^s x 2^s uniform sub-quads of the
 * element, the point evaluation of ext_st (bem_staela3d.f90:654-969, bem_harela3d.f90:702-1048)
 * without the Telles stretch.  s = 0 is the precalculated-dataset rule of ext_pre
 * (bem_staela3d.f90:592-650).  Used to label and to judge rule choices.
 * ---------------------------------------------------------------------------------------------- */
static void ext_rule(const kpar_t *k, const double *X, const double *x_i, int gln, int s,
                     acc_t *a) {
  const int m = 1 << s;
  const double hw = 1.0 / (double)m;
  for (int i1 = 0; i1 < m; i1++)
    for (int i2 = 0; i2 < m; i2++) {
      const double c1 = -1.0 + (2.0 * i1 + 1.0) * hw, c2 = -1.0 + (2.0 * i2 + 1.0) * hw;
      for (int k1 = 0; k1 < gln; k1++)
        for (int k2 = 0; k2 < gln; k2++) {
          double xi[2] = {c1 + hw * GL11X[gln][k1], c2 + hw * GL11X[gln][k2]};
          double x[3], T1[3], T2[3], N[3], n[3], p[9], pjw[9];
          q9_geom(X, xi, x, T1, T2);
          cross(T1, T2, N);
          double jg = sqrt(dot3(N, N));
          for (int c = 0; c < 3; c++) n[c] = N[c] / jg;
          double jw = jg * GL11W[gln][k1] * GL11W[gln][k2] * hw * hw;
          phi_q9(xi, p);
          for (int q = 0; q < 9; q++) pjw[q] = p[q] * jw;
          ext_point(k, x, n, x_i, pjw, pjw, a);
        }
    }
}

int orc_pair_rule_sta(const double *x, const double *x_i, double mu, double nu, int reverse,
                      int gln, int s, double *h, double *g) {
  if (gln < 1 || gln > 32 || s < 0 || s > 8) return -1;
  kpar_t k;
  acc_t a;
  gl_init();
  kpar_static(mu, nu, &k);
  acc_zero(&a);
  ext_rule(&k, x, x_i, gln, s, &a);
  acc_finish(&k, reverse, &a);
  memcpy(h, a.hr, sizeof a.hr);
  memcpy(g, a.gr, sizeof a.gr);
  return 0;
}

int orc_pair_rule_har(const double *x, const double *x_i, double lam_r, double lam_i, double mu_r,
                      double mu_i, double rho, double omega, int reverse, int gln, int s,
                      double *h, double *g) {
  if (gln < 1 || gln > 32 || s < 0 || s > 8) return -1;
  kpar_t k;
  acc_t a;
  gl_init();
  kpar_harmonic(lam_r + _Complex_I * lam_i, mu_r + _Complex_I * mu_i, rho, omega, &k);
  acc_zero(&a);
  ext_rule(&k, x, x_i, gln, s, &a);
  acc_finish(&k, reverse, &a);
  for (int i = 0; i < 81; i++) {
    h[2 * i] = creal(a.hc[i]);
    h[2 * i + 1] = cimag(a.hc[i]);
    g[2 * i] = creal(a.gc[i]);
    g[2 * i + 1] = cimag(a.gc[i]);
  }
  return 0;
}

/* ------------------------------------------------------------------------------------------------
 * Mantic free term (bem_staela3d.f90:228-405; identical in bem_harela3d.f90:365-542 with complex nu
 * whose imaginary part is zero for the reference's visco-elastic regions)
 * ---------------------------------------------------------------------------------------------- */
void orc_freeterm(int m, const double *n, const double *t, double tol, double nu, double *C) {
  double ltol = (tol < 1e-12 || tol > 1e-3) ? 1e-6 : tol;
  double *ln = malloc(sizeof(double) * 3 * (m + 2)), *lt = malloc(sizeof(double) * 3 * (m + 2));
  double *lti = malloc(sizeof(double) * 3 * m), *theta = malloc(sizeof(double) * m);
  int *tc = malloc(sizeof(int) * m);
#define LN(c, i) ln[3 * (i) + (c)]
#define LT(c, i) lt[3 * (i) + (c)]
  for (int i = 1; i <= m; i++)
    for (int c = 0; c < 3; c++) {
      LN(c, i) = n[3 * (i - 1) + c];
      LT(c, i) = t[3 * (i - 1) + c];
    }
  for (int ki = 1; ki <= m - 1; ki++) {
    double e1[3], e2[3], e3[3];
    for (int c = 0; c < 3; c++) { e1[c] = LT(c, ki); e3[c] = LN(c, ki); }
    e2[0] = e3[1] * e1[2] - e3[2] * e1[1];
    e2[1] = e3[2] * e1[0] - e3[0] * e1[2];
    e2[2] = e3[0] * e1[1] - e3[1] * e1[0];
    for (int kj = 1; kj <= m; kj++) {
      lti[3 * (kj - 1) + 0] = e1[0] * LT(0, kj) + e1[1] * LT(1, kj) + e1[2] * LT(2, kj);
      lti[3 * (kj - 1) + 1] = e2[0] * LT(0, kj) + e2[1] * LT(1, kj) + e2[2] * LT(2, kj);
      lti[3 * (kj - 1) + 2] = e3[0] * LT(0, kj) + e3[1] * LT(1, kj) + e3[2] * LT(2, kj);
    }
    int nt = 0;
    for (int kj = ki + 1; kj <= m; kj++)
      if (fabs(lti[3 * (kj - 1) + 2]) <= ltol) tc[nt++] = kj;
    if (nt == 0) {
      /* the reference stops: 'the normals/tangents configuration is not valid' */
      for (int i = 0; i < 9; i++) C[i] = NAN;
      free(ln); free(lt); free(lti); free(theta); free(tc);
      return;
    }
    for (int j = 0; j < nt; j++) {
      int kj = tc[j];
      double a = atan2(lti[3 * (kj - 1) + 1], lti[3 * (kj - 1) + 0]);
      if (a < 0.0) a += 2.0 * C_PI;
      theta[kj - 1] = a;
    }
    double mint = theta[tc[0] - 1];
    int minkj = tc[0];
    for (int j = 1; j < nt; j++)
      if (theta[tc[j] - 1] < mint) { mint = theta[tc[j] - 1]; minkj = tc[j]; }
    for (int c = 0; c < 3; c++) {
      double tmp = LT(c, ki + 1); LT(c, ki + 1) = LT(c, minkj); LT(c, minkj) = tmp;
      tmp = LN(c, ki + 1); LN(c, ki + 1) = LN(c, minkj); LN(c, minkj) = tmp;
    }
  }
  for (int c = 0; c < 3; c++) {
    LN(c, 0) = LN(c, m); LT(c, 0) = LT(c, m);
    LN(c, m + 1) = LN(c, 1); LT(c, m + 1) = LT(c, 1);
  }
  double sum_a = 0;
  for (int ki = 1; ki <= m; ki++) {
    double nx[3];
    nx[0] = LN(1, ki - 1) * LN(2, ki) - LN(2, ki - 1) * LN(1, ki);
    nx[1] = LN(2, ki - 1) * LN(0, ki) - LN(0, ki - 1) * LN(2, ki);
    nx[2] = LN(0, ki - 1) * LN(1, ki) - LN(1, ki - 1) * LN(0, ki);
    double s = nx[0] * LT(0, ki) + nx[1] * LT(1, ki) + nx[2] * LT(2, ki);
    if (s < 0.0) s = -1.0;
    if (s > 0.0) s = 1.0;
    double ndn = LN(0, ki - 1) * LN(0, ki) + LN(1, ki - 1) * LN(1, ki) + LN(2, ki - 1) * LN(2, ki);
    if (ndn > 1.0) ndn = 1.0;
    sum_a = sum_a + s * acos(ndn);
  }
  double cp = 1.0 / (4.0 * C_PI) * (2.0 * C_PI + sum_a);
  double sb[9] = {0};
  for (int ki = 1; ki <= m; ki++) {
    double rr[3], rx[3];
    for (int c = 0; c < 3; c++) rr[c] = LT(c, ki + 1) - LT(c, ki);
    rx[0] = rr[1] * LN(2, ki) - rr[2] * LN(1, ki);
    rx[1] = rr[2] * LN(0, ki) - rr[0] * LN(2, ki);
    rx[2] = rr[0] * LN(1, ki) - rr[1] * LN(0, ki);
    for (int c1 = 0; c1 < 3; c1++)
      for (int c2 = 0; c2 < 3; c2++) sb[c1 + 3 * c2] = sb[c1 + 3 * c2] + rx[c1] * LN(c2, ki);
  }
  for (int c1 = 0; c1 < 3; c1++)
    for (int c2 = 0; c2 < 3; c2++) {
      C[c1 + 3 * c2] = -1.0 / (8.0 * C_PI * (1.0 - nu)) * sb[c1 + 3 * c2];
      if (c1 == c2) C[c1 + 3 * c2] += cp;
    }
#undef LN
#undef LT
  free(ln); free(lt); free(lti); free(theta); free(tc);
}

/* ------------------------------------------------------------------------------------------------
 * Whole-boundary assembly (src/build_lse_mechanics_bem_staela.f90:22-655,659-1305 and the harmonic
 * counterpart; scatter rule of src/assemble_bem_{staela,harela}_equation.f90, BE ordinary case)
 * ---------------------------------------------------------------------------------------------- */
static void utan(const double *X, const double *xi, int which, double *t) {
  double x[3], T1[3], T2[3];
  q9_geom(X, xi, x, T1, T2);
  const double *T = which == 1 ? T1 : T2;
  double nrm = sqrt(T[0] * T[0] + T[1] * T[1] + T[2] * T[2]);
  for (int c = 0; c < 3; c++) t[c] = T[c] / nrm;
}

static void unormal(const double *X, const double *xi, double *n) {
  double x[3], T1[3], T2[3], N[3];
  q9_geom(X, xi, x, T1, T2);
  cross(T1, T2, N);
  double j = sqrt(dot3(N, N));
  for (int c = 0; c < 3; c++) n[c] = N[c] / j;
}

/* boundary tangent tbp at quad9 local node kn (geometry.f90:657-870) */
static void tbp_q9(const double *X, int kn, double *t) {
  double s = 1.0;
  int which = 1;
  switch (kn) {
    case 0: case 4: which = 1; s = 1.0; break;
    case 1: case 5: which = 2; s = 1.0; break;
    case 2: case 6: which = 1; s = -1.0; break;
    case 3: case 7: which = 2; s = -1.0; break;
    default: t[0] = t[1] = t[2] = 0; return;
  }
  utan(X, XI_Q9[kn], which, t);
  for (int c = 0; c < 3; c++) t[c] = s * t[c];
}

static void gather_X(const double *xn, const int *conn, int e, double *X) {
  for (int k = 0; k < 9; k++)
    for (int c = 0; c < 3; c++) X[3 * k + c] = xn[3 * conn[9 * e + k] + c];
}

/* Collocation points of the reference's default formulation (src/assign_default_bem_formulation.f90:
 * 74-93): nodal SBIE, except nodes on an element edge that no other element shares (the boundary of
 * the boundary, lib/fbem/src/data_structures.f90:1620-1640), which are collocated by SBIE-MCA at
 * xi_node (1 - delta) of every element holding them (fbem_move_xi1xi2_from_edge,
 * geometry.f90:5631-5658).  Points node by node; an MCA node's points in element order.  Returns
 * the number of points; xc (3 each), crow, cel (-1 nodal), cxi (2 each) may be NULL. */
int orc_collocation(int nn, int ne, const int *conn, const double *xn, double delta, double *xc,
                    int *crow, int *cel, double *cxi, unsigned char *is_mca) {
  unsigned char *mca = calloc(nn, 1);
  for (int e = 0; e < ne; e++)
    for (int k = 0; k < 4; k++) {
      int a = conn[9 * e + EDGE_Q9[k][0]], b = conn[9 * e + EDGE_Q9[k][1]], c = conn[9 * e + EDGE_Q9[k][2]];
      int shared = 0;
      for (int e2 = 0; e2 < ne && !shared; e2++) {
        if (e2 == e) continue;
        for (int k2 = 0; k2 < 4; k2++) {
          int a2 = conn[9 * e2 + EDGE_Q9[k2][0]], b2 = conn[9 * e2 + EDGE_Q9[k2][1]];
          int c2 = conn[9 * e2 + EDGE_Q9[k2][2]];
          if (c2 == c && ((a2 == a && b2 == b) || (a2 == b && b2 == a))) { shared = 1; break; }
        }
      }
      if (!shared) mca[a] = mca[b] = mca[c] = 1;
    }
  int n = 0;
  for (int i = 0; i < nn; i++) {
    if (!mca[i]) {
      if (xc) for (int c = 0; c < 3; c++) xc[3 * n + c] = xn[3 * i + c];
      if (crow) crow[n] = i;
      if (cel) cel[n] = -1;
      if (cxi) cxi[2 * n] = cxi[2 * n + 1] = 0.0;
      n++;
      continue;
    }
    for (int e = 0; e < ne; e++)
      for (int k = 0; k < 9; k++) {
        if (conn[9 * e + k] != i) continue;
        double xi[2] = {XI_Q9[k][0] * (1.0 - delta), XI_Q9[k][1] * (1.0 - delta)}, X[27], p[9];
        gather_X(xn, conn, e, X);
        phi_q9(xi, p);
        if (xc)
          for (int c = 0; c < 3; c++) {
            double v = 0.0;
            for (int m = 0; m < 9; m++) v += p[m] * X[3 * m + c];
            xc[3 * n + c] = v;
          }
        if (crow) crow[n] = i;
        if (cel) cel[n] = e;
        if (cxi) { cxi[2 * n] = xi[0]; cxi[2 * n + 1] = xi[1]; }
        n++;
      }
  }
  if (is_mca) memcpy(is_mca, mca, nn);
  free(mca);
  return n;
}

static int assemble(const kpar_t *k, int nn, int ne, const double *xn, const int *conn,
                    const int *ctype, const double *cval, double nu, double qsi_re, int ns,
                    int n_ps, const int *ps_gln, double tol, int node_begin, int node_end, double *A,
                    double *b, long long *stats) {
  gl_init();
  /* rows of the collocation nodes [node_begin, node_end) only: A is (3 (node_end - node_begin)) x n,
   * column-major, b has as many rows (the whole system when the range is every node) */
  const long long r0 = 3LL * node_begin, ldr = 3LL * (node_end - node_begin);
  const int harm = k->harmonic;
  zc *Ac = (zc *)A, *bc = (zc *)b;
  /* node normals: normalised sum of the element normals at the node */
  double *nfn = calloc(3 * (size_t)nn, sizeof(double));
  int *first_e = malloc(sizeof(int) * nn), *first_k = malloc(sizeof(int) * nn);
  for (int i = 0; i < nn; i++) first_e[i] = -1;
  for (int e = 0; e < ne; e++) {
    double X[27];
    gather_X(xn, conn, e, X);
    for (int kn = 0; kn < 9; kn++) {
      int j = conn[9 * e + kn];
      double nv[3];
      unormal(X, XI_Q9[kn], nv);
      for (int c = 0; c < 3; c++) nfn[3 * j + c] += nv[c];
      if (first_e[j] < 0) { first_e[j] = e; first_k[j] = kn; }
    }
  }
  for (int i = 0; i < nn; i++) {
    double s = sqrt(dot3(nfn + 3 * i, nfn + 3 * i));
    for (int c = 0; c < 3; c++) nfn[3 * i + c] /= s;
  }
  long long st[6] = {0};
  qsp_t qsp;
  qsp_setup(qsi_re, &qsp);
  /* collocation points: nodal, and SBIE-MCA (delta 0.05, default_mca_boundary_delta) on open edges */
  const int ncp = orc_collocation(nn, ne, conn, xn, 0.05, NULL, NULL, NULL, NULL, NULL);
  double *xc = malloc(sizeof(double) * 3 * ncp), *cxi = malloc(sizeof(double) * 2 * ncp);
  int *crow = malloc(sizeof(int) * ncp), *cel = malloc(sizeof(int) * ncp);
  unsigned char *is_mca = malloc(nn);
  orc_collocation(nn, ne, conn, xn, 0.05, xc, crow, cel, cxi, is_mca);
  /* first collocation point of every node */
  int *cp0 = malloc(sizeof(int) * (nn + 1));
  for (int i = 0; i <= nn; i++) cp0[i] = ncp;
  for (int q = ncp - 1; q >= 0; q--) cp0[crow[q]] = q;
  /* collocation order of the reference: nodes in order of first appearance in the element list */
  int *order = malloc(sizeof(int) * nn), no = 0;
  {
    char *seen = calloc(nn, 1);
    for (int e = 0; e < ne; e++)
      for (int kn = 0; kn < 9; kn++) {
        int j = conn[9 * e + kn];
        if (!seen[j]) { seen[j] = 1; order[no++] = j; }
      }
    free(seen);
  }
#pragma omp parallel
  {
    long long lst[6] = {0};
#pragma omp for schedule(dynamic)
    for (int e = 0; e < ne; e++) {
      elem_t el;
      double X[27];
      gather_X(xn, conn, e, X);
      elem_init(&el, X, qsi_re, n_ps, ps_gln);
      const int *en = conn + 9 * e;
      for (int oi = 0; oi < no; oi++) {
        int i = order[oi];
        if (i < node_begin || i >= node_end) continue;
        for (int q = cp0[i]; q < ncp && crow[q] == i; q++) {
        acc_t a;
        trace_t tr;
        sbie_auto(k, &el, 0, xc + 3 * q, &qsp, ns, &a, &tr);
        lst[0]++;
        if (tr.delta) lst[1]++;
        else if (tr.ps > 0) lst[2]++;
        else { lst[3]++; lst[5] += tr.leaves; }
        lst[4] += a.npts;
#pragma omp critical(orc_scatter)
        for (int il = 0; il < 3; il++) {
          long long row = 3LL * i + il;
          for (int ik = 0; ik < 3; ik++)
            for (int kk = 0; kk < 9; kk++) {
              int sn = en[kk];
              long long col = 3LL * sn + ik;
              int o = kk + 9 * il + 27 * ik;
              int ct = ctype[3 * sn + ik];
              double cv = cval[3 * sn + ik];
              if (!harm) {
                if (ct == 0) {
                  A[(row - r0) + ldr * col] -= a.gr[o];
                  b[row - r0] -= a.hr[o] * cv;
                } else if (ct == 1) {
                  A[(row - r0) + ldr * col] += a.hr[o];
                  b[row - r0] += a.gr[o] * cv;
                } else {
                  A[(row - r0) + ldr * col] += a.hr[o];
                  b[row - r0] += a.gr[o] * cv * nfn[3 * sn + ik];
                }
              } else {
                if (ct == 0) {
                  Ac[(row - r0) + ldr * col] -= a.gc[o];
                  bc[row - r0] -= a.hc[o] * cv;
                } else if (ct == 1) {
                  Ac[(row - r0) + ldr * col] += a.hc[o];
                  bc[row - r0] += a.gc[o] * cv;
                } else {
                  Ac[(row - r0) + ldr * col] += a.hc[o];
                  bc[row - r0] += a.gc[o] * cv * nfn[3 * sn + ik];
                }
              }
            }
        }
        }
      }
      elem_free(&el);
    }
#pragma omp critical(orc_stats)
    for (int s = 0; s < 6; s++) st[s] += lst[s];
  }
  /* free terms, added once per node at its first element (build_lse_mechanics_bem_staela.f90:210-549) */
  for (int i = node_begin; i < node_end; i++) {
    int e = first_e[i], kn = first_k[i];
    if (e < 0) continue;
    if (is_mca[i]) {
      /* SBIE-MCA: (1/2) phi(xi_i) I on the nodes of each element holding a point of the node
       * (src/build_lse_mechanics_bem_staela.f90:444-457) */
      for (int q = cp0[i]; q < ncp && crow[q] == i; q++) {
        double p[9];
        phi_q9(cxi + 2 * q, p);
        for (int m = 0; m < 9; m++) {
          int sn = conn[9 * cel[q] + m];
          for (int il = 0; il < 3; il++) {
            long long row = 3LL * i + il, col = 3LL * sn + il;
            double hv = 0.5 * p[m];
            int ct = ctype[3 * sn + il];
            double uv = cval[3 * sn + il];
            if (!harm) {
              if (ct == 0) b[row - r0] -= hv * uv;
              else A[(row - r0) + ldr * col] += hv;
            } else {
              if (ct == 0) bc[row - r0] -= hv * uv;
              else Ac[(row - r0) + ldr * col] += hv;
            }
          }
        }
      }
      continue;
    }
    double C[9];
    if (kn < 8) {
      int m = 0;
      for (int e2 = 0; e2 < ne; e2++)
        for (int k2 = 0; k2 < 9; k2++)
          if (conn[9 * e2 + k2] == i) m++;
      double *ns_ = malloc(sizeof(double) * 3 * m), *ts = malloc(sizeof(double) * 3 * m);
      m = 0;
      for (int e2 = 0; e2 < ne; e2++)
        for (int k2 = 0; k2 < 9; k2++)
          if (conn[9 * e2 + k2] == i) {
            double X[27];
            gather_X(xn, conn, e2, X);
            unormal(X, XI_Q9[k2], ns_ + 3 * m);
            tbp_q9(X, k2, ts + 3 * m);
            m++;
          }
      orc_freeterm(m, ns_, ts, tol, nu, C);
      free(ns_);
      free(ts);
    } else {
      for (int c = 0; c < 9; c++) C[c] = (c % 4 == 0) ? 0.5 : 0.0;
    }
    for (int il = 0; il < 3; il++) {
      long long row = 3LL * i + il;
      for (int ik = 0; ik < 3; ik++) {
        long long col = 3LL * i + ik;
        double cv = C[il + 3 * ik];
        int ct = ctype[3 * i + ik];
        double uv = cval[3 * i + ik];
        if (!harm) {
          if (ct == 0) b[row - r0] -= cv * uv;
          else A[(row - r0) + ldr * col] += cv;
        } else {
          if (ct == 0) bc[row - r0] -= cv * uv;
          else Ac[(row - r0) + ldr * col] += cv;
        }
      }
    }
  }
  free(order);
  free(nfn);
  free(first_e);
  free(first_k);
  free(xc);
  free(cxi);
  free(crow);
  free(cel);
  free(is_mca);
  free(cp0);
  if (stats)
    for (int s = 0; s < 6; s++) stats[s] = st[s];
  return 0;
}

int orc_assemble_sta(int nn, int ne, const double *xn, const int *conn, const int *ctype,
                     const double *cval, double mu, double nu, double qsi_re, int ns, int n_ps,
                     const int *ps_gln, double tol, int node_begin, int node_end, double *A,
                     double *b, long long *stats) {
  kpar_t k;
  kpar_static(mu, nu, &k);
  return assemble(&k, nn, ne, xn, conn, ctype, cval, nu, qsi_re, ns, n_ps, ps_gln, tol, node_begin,
                  node_end, A, b, stats);
}

int orc_assemble_har(int nn, int ne, const double *xn, const int *conn, const int *ctype,
                     const double *cval, double lam_r, double lam_i, double mu_r, double mu_i,
                     double rho, double omega, double nu, double qsi_re, int ns, int n_ps,
                     const int *ps_gln, double tol, int node_begin, int node_end, double *A,
                     double *b, long long *stats) {
  kpar_t k;
  kpar_harmonic(lam_r + _Complex_I * lam_i, mu_r + _Complex_I * mu_i, rho, omega, &k);
  return assemble(&k, nn, ne, xn, conn, ctype, cval, nu, qsi_re, ns, n_ps, ps_gln, tol, node_begin,
                  node_end, A, b, stats);
}

/* Integration class of every pair (e, i), coded as the GPU library codes it: 0..n_ps-1 the
 * precalculated dataset index, n_ps adaptive, n_ps+1 singular (decision sequence of
 * bem_staela3d.f90:549-587). out has ne*nn bytes, index e*nn + i. */
void orc_pair_classes_pts(int npt, const double *xp, int ne, const double *xn, const int *conn,
                          double qsi_re, int n_ps, const int *ps_gln, unsigned char *out);

void orc_pair_classes(int nn, int ne, const double *xn, const int *conn, double qsi_re, int n_ps,
                      const int *ps_gln, unsigned char *out) {
  orc_pair_classes_pts(nn, xn, ne, xn, conn, qsi_re, n_ps, ps_gln, out);
}

/* the same for arbitrary points xp (npt x 3): out has ne*npt bytes, index e*npt + p */
void orc_pair_classes_pts(int npt, const double *xp, int ne, const double *xn, const int *conn,
                          double qsi_re, int n_ps, const int *ps_gln, unsigned char *out) {
  const int nn = npt;
  gl_init();
  qsp_t q;
  qsp_setup(qsi_re, &q);
  int gmax = ps_gln[n_ps - 1];
#pragma omp parallel for schedule(dynamic)
  for (int e = 0; e < ne; e++) {
    double X[27], cl, centre[3], radius;
    int gln_far;
    gather_X(xn, conn, e, X);
    orc_element_data(X, qsi_re, &cl, &gln_far, centre, &radius);
    for (int i = 0; i < nn; i++) {
      const double *x_i = xp + 3 * i;
      double r[3] = {centre[0] - x_i[0], centre[1] - x_i[1], centre[2] - x_i[2]};
      double rmin = sqrt(dot3(r, r)) - radius, barxi[2] = {0, 0}, d;
      int c;
      if (rmin > 4.0 * radius) {
        d = rmin / cl;
      } else {
        int method;
        orc_nearest_quad9(X, cl, x_i, barxi, &rmin, &d, &method);
      }
      if (rmin <= 4.0 * radius && d <= 1e-12) {
        c = n_ps + 1;
      } else {
        int gln_near = qs_estimate(&q, 0, 2, 5, d, barxi);
        int gln = gln_far > gln_near ? gln_far : gln_near;
        c = n_ps;
        if (gln <= gmax && gln_near > 0)
          for (int k = 0; k < n_ps; k++)
            if (ps_gln[k] >= gln) { c = k; break; }
      }
      out[(long long)e * nn + i] = (unsigned char)c;
    }
  }
}

/* ------------------------------------------------------------------------------------------------
 * Interior points (src/calculate_internal_points_mechanics_static.f90:76-89 and
 * src/calculate_internal_points_mechanics_bem_harela.f90:160-210, 355-363): for every element and
 * point, sbie_auto at the point, u_l += g(n,l,k) t_k(node n) - h(n,l,k) u_k(node n); a point on a
 * boundary node (distance / characteristic length of the node's first element < 1e-12) takes the
 * node's displacement.  ub, tb: (nn x 3), out: (npt x 3), complex interleaved when harmonic.
 * ---------------------------------------------------------------------------------------------- */
static int points(const kpar_t *k, int nn, int ne, const double *xn, const int *conn, int npt,
                  const double *xp, double qsi_re, int ns, int n_ps, const int *ps_gln,
                  const double *ub, const double *tb, double *out) {
  gl_init();
  const int harm = k->harmonic;
  const zc *ubc = (const zc *)ub, *tbc = (const zc *)tb;
  zc *oc = (zc *)out;
  qsp_t qsp;
  qsp_setup(qsi_re, &qsp);
  if (harm) for (long long q = 0; q < 3LL * npt; q++) oc[q] = 0;
  else for (long long q = 0; q < 3LL * npt; q++) out[q] = 0.0;
#pragma omp parallel
  {
    double *acc_r = harm ? NULL : calloc(3 * (size_t)npt, sizeof(double));
    zc *acc_c = harm ? calloc(3 * (size_t)npt, sizeof(zc)) : NULL;
#pragma omp for schedule(dynamic)
    for (int e = 0; e < ne; e++) {
      elem_t el;
      double X[27];
      gather_X(xn, conn, e, X);
      elem_init(&el, X, qsi_re, n_ps, ps_gln);
      const int *en = conn + 9 * e;
      for (int p = 0; p < npt; p++) {
        acc_t a;
        trace_t tr;
        sbie_auto(k, &el, 0, xp + 3 * p, &qsp, ns, &a, &tr);
        for (int il = 0; il < 3; il++)
          for (int ik = 0; ik < 3; ik++)
            for (int kk = 0; kk < 9; kk++) {
              int sn = en[kk], o = kk + 9 * il + 27 * ik;
              if (harm) acc_c[3 * p + il] += a.gc[o] * tbc[3 * sn + ik] - a.hc[o] * ubc[3 * sn + ik];
              else acc_r[3 * p + il] += a.gr[o] * tb[3 * sn + ik] - a.hr[o] * ub[3 * sn + ik];
            }
      }
      elem_free(&el);
    }
#pragma omp critical(orc_points)
    for (long long q = 0; q < 3LL * npt; q++) {
      if (harm) oc[q] += acc_c[q];
      else out[q] += acc_r[q];
    }
    free(acc_r);
    free(acc_c);
  }
  /* points on boundary nodes */
  int *first_e = malloc(sizeof(int) * nn);
  for (int i = 0; i < nn; i++) first_e[i] = -1;
  for (int e = ne - 1; e >= 0; e--)
    for (int kk = 0; kk < 9; kk++) first_e[conn[9 * e + kk]] = e;
  for (int p = 0; p < npt; p++)
    for (int i = 0; i < nn; i++) {
      if (first_e[i] < 0) continue;
      double X[27], cl, cen[3], rad;
      int nphi;
      gather_X(xn, conn, first_e[i], X);
      orc_element_data(X, qsi_re, &cl, &nphi, cen, &rad);
      double d[3] = {xp[3 * p] - xn[3 * i], xp[3 * p + 1] - xn[3 * i + 1], xp[3 * p + 2] - xn[3 * i + 2]};
      if (sqrt(dot3(d, d)) / cl < 1e-12)
        for (int c = 0; c < 3; c++) {
          if (harm) oc[3 * p + c] = ubc[3 * i + c];
          else out[3 * p + c] = ub[3 * i + c];
        }
    }
  free(first_e);
  return 0;
}

int orc_points_sta(int nn, int ne, const double *xn, const int *conn, int npt, const double *xp,
                   double mu, double nu, double qsi_re, int ns, int n_ps, const int *ps_gln,
                   const double *ub, const double *tb, double *out) {
  kpar_t k;
  kpar_static(mu, nu, &k);
  return points(&k, nn, ne, xn, conn, npt, xp, qsi_re, ns, n_ps, ps_gln, ub, tb, out);
}

int orc_points_har(int nn, int ne, const double *xn, const int *conn, int npt, const double *xp,
                   double lam_r, double lam_i, double mu_r, double mu_i, double rho, double omega,
                   double qsi_re, int ns, int n_ps, const int *ps_gln, const double *ub,
                   const double *tb, double *out) {
  kpar_t k;
  kpar_harmonic(lam_r + _Complex_I * lam_i, mu_r + _Complex_I * mu_i, rho, omega, &k);
  return points(&k, nn, ne, xn, conn, npt, xp, qsi_re, ns, n_ps, ps_gln, ub, tb, out);
}
