// BEM influence-matrix assembly on MI355X (gfx950), FP64.
//
// Pipeline of one mesh (all kernels on one stream):
//   mesh creation (geometry only, once):
//     k_elem_setup     per element: characteristic length, gln_far, bounding ball
//     k_precalc        per element x precalculated point: x, n, phi*J*w for Gauss-Legendre 2..9
//     k_node_normals   per node: normalised sum of element normals (ctype 10 loads)
//     k_classify       per element block x collocation nodes: integration class of every pair
//     k_fill_lists     pair ids grouped by class (precalculated rule r, adaptive, singular),
//                      each class list ordered by element so that batches share element data
//   every assembly (material / frequency dependent):
//     k_zero           A, b := 0
//     k_regular_tile<H> one workgroup per tile of up to 14 pairs of one element and one
//                      precalculated rule, one launch per element colour: threads evaluate the
//                      kernel at (pair, point) items into LDS, then each thread contracts one
//                      (pair, h|g, component) row against the element's 9 weights and adds the
//                      row to A (or b)
//     k_regular<H,MODE> per-rule variant writing per-pair blocks or interior displacements
//     k_adaptive<H>    one wave per quasi-singular pair: lane 0 walks the reference's 4-way
//                      subdivision tree, all lanes integrate the Telles leaves
//     k_singular<H>    one wave per pair with the collocation node on the element: polar
//                      sub-triangles, CPV regularisation, Stokes line integrals on the edges
//     k_freeterm<H>    per node: Mantic free term (or I/2 at element-interior nodes)
// Scatter follows the BE-ordinary rule of src/assemble_bem_{staela,harela}_equation.f90: the
// unknown of node j, direction k is column 3j+k; u-known directions move h to b and put -g in A.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "mfbem.h"
#include "mfbem_device.h"
#include "mfbem_internal.h"

namespace mfb {

__constant__ GLTables c_gl;
__device__ inline const GLTables& gl() { return c_gl; }

// Polar sub-triangles of a collocation point on one of the 9 nodes of a quad9 (nodal collocation:
// every singular pair has its point there).  Computed once on the host, whose libm is the one the
// reference calls: the number of angular points, 5 + nint(25 dtheta / (pi/2)), sits exactly on a
// half-integer at corner nodes (dtheta = pi/4) and a 1-ulp difference in asin flips it.
struct PolarSet {
  int n;
  int st[8];
  int ngt[8];
  double th[16];
  double thp[16];
};
__constant__ PolarSet c_polar[9];

constexpr int kWave = 64;
// adaptive subdivision: leaves integrated per batch, and the walk's stack (4 entries pushed per
// level: at most 3 ns_max + 1 live entries, so ns_max <= kMaxNs)
constexpr int kMaxLeaves = 128;
constexpr int kStack = 64;
constexpr int kMaxNs = 20;
// largest precalculated rule: g^2 points x 360 B of kernel values and weights (harmonic) within
// the 160 KB of LDS a gfx950 workgroup may hold
constexpr int kMaxPrecalGln = 21;
constexpr size_t kMaxLds = 160 * 1024;
// quadrature points staged in LDS per pass by the adaptive and singular kernels: one per lane
// (their LDS stays under 40 KB)
constexpr int kChunk = 64;
constexpr int kMaxIncid = 24;

// Collocation points: the reference's nodal SBIE puts one point on every node; a node on the open
// edge of a boundary (an element edge with no neighbour) is collocated by SBIE-MCA instead, with
// one point inside every element that holds the node, all adding to the node's equations
// (src/assign_default_bem_formulation.f90:74-93, src/build_lse_mechanics_bem_staela.f90:890-925).
// Pairs are (collocation point cp, element e), pair id e * ncp + cp.
struct DevMesh {
  int nn, ne, n_ps, npts;  // npts: precalculated points per element (sum of gln^2)
  int ncp;                 // collocation points
  int ps_gln[kMaxPrecalsets], ps_off[kMaxPrecalsets + 1];
  int ns_max;
  double qsi_re, tol;
  const double* x;
  const int* conn;
  const int* ctype;
  const double* cval;
  const double* elem;  // 8 per element: cl, radius, centre(3), gln_far, 0, 0
  const double* ps;    // (ne, npts, kPsStride)
  const double* nfn;   // (nn, 3)
  const int* inc_ptr;  // node -> (element, local node) incidences, ordered by element
  const int* inc_e;
  const int* inc_k;
  const double* xc;             // (ncp, 3) collocation point coordinates
  const int* crow;              // (ncp) node whose equations (rows 3 crow + l) the point adds to
  const unsigned char* cshared; // (ncp) 1 when other points add to the same rows (MCA)
  const int* cel;               // (ncp) element holding an MCA point, -1 for a nodal point
  const double* cxi;            // (ncp, 2) local coordinates of an MCA point in cel
  const unsigned char* nmca;    // (nn) 1 for a node collocated by MCA points
};

// ------------------------------------------------------------------------------------------------
// small helpers
// ------------------------------------------------------------------------------------------------
template <bool H>
struct Val;
template <>
struct Val<false> {
  typedef double T;
};
template <>
struct Val<true> {
  typedef zd T;
};

__device__ inline void atomic_add(double* a, double v) { unsafeAtomicAdd(a, v); }
__device__ inline void atomic_add(zd* a, zd v) {
  unsafeAtomicAdd(&a->re, v.re);
  unsafeAtomicAdd(&a->im, v.im);
}
__device__ inline double mag(double a) { return fabs(a); }
__device__ inline double mag(zd a) { return hypot(a.re, a.im); }
__device__ inline double zero_of(double) { return 0.0; }
__device__ inline zd zero_of(zd) { return mkz(0.0); }
template <typename T>
__device__ inline T from_real(double v);
template <>
__device__ inline double from_real<double>(double v) { return v; }
template <>
__device__ inline zd from_real<zd>(double v) { return mkz(v); }
__device__ inline double scale_h(const KernelParams& k, double v) { return v * k.ctet1; }
__device__ inline double scale_g(const KernelParams& k, double v) { return v * k.cteu1; }
__device__ inline zd scale_h(const KernelParams& k, zd v) { return mkz(k.h.cte_t[0], k.h.cte_t[1]) * v; }
__device__ inline zd scale_g(const KernelParams& k, zd v) { return mkz(k.h.cte_u[0], k.h.cte_u[1]) * v; }

// Exterior kernel values at one point before the constants (bem_staela3d.f90:636-645;
// bem_harela3d.f90:657-690) and the regularised values of the singular integration
// (bem_staela3d.f90:1303-1322; bem_harela3d.f90:1418-1438):
// h += ft*phi*jw + fc*(phi-phi_i)*jw, g += fu*phi*jw.
// Same kernels, written component by component into an LDS row (keeps the 18-27 values out of
// registers): row[c] = ft, row[9 + c] = fu, row[18 + c] = fc (entries `stride` apart).
__device__ inline void ext_kernel_row(const KernelParams& k, const double* x, const double* n,
                                      const double* x_i, double* row, int stride = 1) {
  double rv[3] = {x[0] - x_i[0], x[1] - x_i[1], x[2] - x_i[2]};
  double r = sqrt(dot3(rv, rv)), d1r = 1.0 / r, d1r2 = d1r * d1r;
  double dr[3] = {rv[0] * d1r, rv[1] * d1r, rv[2] * d1r};
  double drdn = dot3(dr, n);
  for (int il = 0; il < 3; il++)
    for (int ik = 0; ik < 3; ik++) {
      double dk = il == ik ? 1.0 : 0.0;
      row[stride * (9 + 3 * il + ik)] = d1r * (k.cteu2 * dk + dr[il] * dr[ik]);
      row[stride * (3 * il + ik)] = d1r2 * ((k.ctet2 * dk + 3.0 * dr[il] * dr[ik]) * drdn +
                                            k.ctet2 * (n[il] * dr[ik] - n[ik] * dr[il]));
    }
}

__device__ inline void ext_kernel_row(const KernelParams& k, const double* x, const double* n,
                                      const double* x_i, zd* row, int stride = 1) {
  double rv[3] = {x[0] - x_i[0], x[1] - x_i[1], x[2] - x_i[2]};
  double r = sqrt(dot3(rv, rv)), d1r = 1.0 / r;
  double dr[3] = {rv[0] * d1r, rv[1] * d1r, rv[2] * d1r};
  double drdn = dot3(dr, n);
  zd psi, chi, TT1, TT2, TT3;
  har_factors(k.h, r, true, psi, chi, TT1, TT2, TT3);
  for (int il = 0; il < 3; il++)
    for (int ik = 0; ik < 3; ik++) {
      double dk = il == ik ? 1.0 : 0.0;
      row[stride * (9 + 3 * il + ik)] = psi * dk - chi * dr[il] * dr[ik];
      row[stride * (3 * il + ik)] = TT1 * dr[il] * dr[ik] * drdn +
                                    TT2 * (drdn * dk + dr[ik] * n[il]) + TT3 * dr[il] * n[ik];
    }
}

__device__ inline void sing_kernel_row(const KernelParams& k, const double* x, const double* n,
                                       const double* x_i, double* row) {
  double rv[3] = {x[0] - x_i[0], x[1] - x_i[1], x[2] - x_i[2]};
  double r = sqrt(dot3(rv, rv)), d1r = 1.0 / r, d1r2 = d1r * d1r;
  double dr[3] = {rv[0] * d1r, rv[1] * d1r, rv[2] * d1r};
  double drdn = dot3(dr, n);
  for (int il = 0; il < 3; il++)
    for (int ik = 0; ik < 3; ik++) {
      double dk = il == ik ? 1.0 : 0.0;
      row[9 + 3 * il + ik] = d1r * (k.cteu2 * dk + dr[il] * dr[ik]);
      row[3 * il + ik] = d1r2 * drdn * (k.ctet2 * dk + 3.0 * dr[il] * dr[ik]);
      row[18 + 3 * il + ik] = d1r2 * k.ctet2 * (n[il] * dr[ik] - n[ik] * dr[il]);
    }
}

__device__ inline void sing_kernel_row(const KernelParams& k, const double* x, const double* n,
                                       const double* x_i, zd* row) {
  double rv[3] = {x[0] - x_i[0], x[1] - x_i[1], x[2] - x_i[2]};
  double r = sqrt(dot3(rv, rv)), d1r = 1.0 / r, d1r2 = d1r * d1r;
  double dr[3] = {rv[0] * d1r, rv[1] * d1r, rv[2] * d1r};
  double drdn = dot3(dr, n);
  zd psi, chi, TT1, TT2, TT3;
  har_factors(k.h, r, false, psi, chi, TT1, TT2, TT3);
  zd t1s = mkz(k.h.T1[0][0], k.h.T1[0][1]) * d1r2, t2s = mkz(k.h.T2[0][0], k.h.T2[0][1]) * d1r2;
  for (int il = 0; il < 3; il++)
    for (int ik = 0; ik < 3; ik++) {
      double dk = il == ik ? 1.0 : 0.0;
      row[9 + 3 * il + ik] = psi * dk - chi * dr[il] * dr[ik];
      row[3 * il + ik] = (TT1 + t1s) * dr[il] * dr[ik] * drdn + (TT2 + t2s) * drdn * dk +
                         TT2 * dr[ik] * n[il] + TT3 * dr[il] * n[ik];
      row[18 + 3 * il + ik] = t2s * (n[il] * dr[ik] - n[ik] * dr[il]);
    }
}

// Scatter of one (collocation node i, component il/ik, h|g) row of 9 node values (already
// multiplied by the kernel constants) into A and b: BE ordinary boundary of
// src/assemble_bem_staela_equation.f90:69-98 and src/assemble_bem_harela_equation.f90:78-113.
template <typename T>
__device__ inline void scatter_row(const DevMesh& m, const int* en, int i, int il, int ik,
                                   bool is_g, const T* v, T* A, long long lda, T* b,
                                   bool plain = false) {
  long long row = 3LL * i + il;
  T bs = zero_of(v[0]);
  bool has_b = false;
  for (int n = 0; n < 9; n++) {
    int sn = en[n];
    long long col = 3LL * sn + ik;
    int ct = m.ctype[3 * sn + ik];
    double cv = m.cval[3 * sn + ik];
    if (!is_g) {
      if (ct == 0) { bs = bs - v[n] * cv; has_b = true; }
      else if (plain) A[row + lda * col] += v[n];
      else atomic_add(&A[row + lda * col], v[n]);
    } else {
      if (ct == 0) {
        if (plain) A[row + lda * col] += -v[n];
        else atomic_add(&A[row + lda * col], -v[n]);
      }
      else if (ct == 1) { bs = bs + v[n] * cv; has_b = true; }
      else { bs = bs + v[n] * (cv * m.nfn[3 * sn + ik]); has_b = true; }
    }
  }
  if (has_b) atomic_add(&b[row], bs);
}

// Third output of the integration kernels: the displacement at interior points,
// u_l(x) = sum over pairs (x, e) of g(n, l, k) t_k(node n) - h(n, l, k) u_k(node n)
// (src/calculate_internal_points_mechanics_static.f90:76-89, ..._bem_harela.f90:355-363): the
// pair's 9x3x3 blocks are contracted with the boundary values as they are integrated.
template <typename T>
struct VecOut {
  const T* u;  // (nn, 3) boundary displacements
  const T* t;  // (nn, 3) boundary tractions
  T* out;      // (ncp, 3) interior displacements (accumulated)
};

__device__ inline void lds_add(double* a, double v) { atomicAdd(a, v); }
__device__ inline void lds_add(zd* a, zd v) {
  atomicAdd(&a->re, v.re);
  atomicAdd(&a->im, v.im);
}

// mesh setup and classification decide rules: plain FP64 rounding, as in the geometry helpers
#pragma clang fp contract(off)

// ------------------------------------------------------------------------------------------------
// mesh setup kernels
// ------------------------------------------------------------------------------------------------
__device__ inline void load_X(const DevMesh& m, int e, double* X) {
  for (int k = 0; k < 9; k++) {
    int j = m.conn[9 * e + k];
    for (int c = 0; c < 3; c++) X[3 * k + c] = m.x[3 * j + c];
  }
}

// fbem_qs_phijac_ngp_2d (quasisingular_integration.f90:812-975) for quad9 geometry and functions
__device__ int phijac_ngp_q9(const double* X, double eh) {
  const GLTables& G = gl();
  double leh = eh <= 0 ? 1e-6 : (eh < 1e-12 ? 1e-12 : eh);
  double in[9] = {0}, old[9];
  int j = 1;
  bool cont = true;
  while (cont) {
    for (int l = 0; l < 9; l++) { old[l] = in[l]; in[l] = 0; }
    for (int k1 = 0; k1 < j; k1++)
      for (int k2 = 0; k2 < j; k2++) {
        double d1[9], d2[9], p[9], T1[3] = {0, 0, 0}, T2[3] = {0, 0, 0}, N[3];
        dphi_q9(G.x11[j][k1], G.x11[j][k2], d1, d2);
        for (int l = 0; l < 9; l++)
          for (int c = 0; c < 3; c++) {
            T1[c] += d1[l] * X[3 * l + c];
            T2[c] += d2[l] * X[3 * l + c];
          }
        cross3(T1, T2, N);
        double jw = sqrt(dot3(N, N)) * G.w11[j][k1] * G.w11[j][k2];
        phi_q9(G.x11[j][k1], G.x11[j][k2], p);
        for (int l = 0; l < 9; l++) in[l] += (p[l] + 1.0) * jw;
      }
    if (j == 1) {
      j++;
    } else {
      cont = false;
      for (int l = 0; l < 9; l++)
        if (fabs((old[l] - in[l]) / in[l]) > leh) { cont = true; j++; break; }
      if (j == 33) cont = false;
    }
  }
  return j - 1;
}

// fbem_geometry_element_ball (geometry.f90:3904-4235), quad branch
__device__ void ball_q9(const double* X, int glp, double* centre, double* radius) {
  const GLTables& G = gl();
  double es = 0, xm[3] = {0, 0, 0};
  for (int k1 = 0; k1 < glp; k1++)
    for (int k2 = 0; k2 < glp; k2++) {
      double x[3], T1[3], T2[3], N[3];
      q9_geom(X, G.x11[glp][k1], G.x11[glp][k2], x, T1, T2);
      cross3(T1, T2, N);
      double jw = sqrt(dot3(N, N)) * G.w11[glp][k1] * G.w11[glp][k2];
      es += jw;
      for (int c = 0; c < 3; c++) xm[c] += x[c] * jw;
    }
  for (int c = 0; c < 3; c++) centre[c] = xm[c] / es;
  double r = 0;
  for (int k = 0; k < 9; k++) {
    double v[3] = {X[3 * k] - centre[0], X[3 * k + 1] - centre[1], X[3 * k + 2] - centre[2]};
    double t = sqrt(dot3(v, v));
    if (r < t) r = t;
  }
  for (int k1 = 0; k1 < glp; k1++)
    for (int k2 = 0; k2 < glp; k2++) {
      double x[3];
      q9_x(X, G.x11[glp][k1], G.x11[glp][k2], x);
      double v[3] = {x[0] - centre[0], x[1] - centre[1], x[2] - centre[2]};
      double t = sqrt(dot3(v, v));
      if (r < t) r = t;
    }
  *radius = r;
}

__global__ void k_elem_setup(DevMesh m, double* elem) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  double X[27], c[3], r;
  load_X(m, e, X);
  double cl = clen_q9(X, 1e-9);
  int nphi = phijac_ngp_q9(X, m.qsi_re);
  ball_q9(X, nphi, c, &r);
  // shape descriptors for the learned rule choice: aspect = longest / shortest edge (arc length);
  // curvature = largest |d2x/dxi2| at the centre (x7 - 2 x8 + x5, x4 - 2 x8 + x6) over cl
  double lmin = 1e300, lmax = 0;
  for (int k = 0; k < 4; k++) {
    double ye[9];
    for (int j = 0; j < 3; j++)
      for (int q = 0; q < 3; q++) ye[3 * j + q] = X[3 * kEdgeQ9[k][j] + q];
    const double le = length_l3(ye, 1e-9);
    lmin = fmin(lmin, le);
    lmax = fmax(lmax, le);
  }
  double c1 = 0, c2 = 0;
  for (int q = 0; q < 3; q++) {
    const double a = X[21 + q] - 2.0 * X[24 + q] + X[15 + q];
    const double b = X[12 + q] - 2.0 * X[24 + q] + X[18 + q];
    c1 += a * a;
    c2 += b * b;
  }
  double* o = elem + 8 * (long long)e;
  o[0] = cl; o[1] = r; o[2] = c[0]; o[3] = c[1]; o[4] = c[2]; o[5] = (double)nphi;
  o[6] = lmax / fmax(lmin, 1e-300);
  o[7] = sqrt(fmax(c1, c2)) / cl;
}

// precalculated datasets, quad branch of init_precalculated_datasets (bem_general.f90:462-530)
__global__ void k_precalc(DevMesh m, double* ps) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)m.ne * m.npts) return;
  int e = (int)(t / m.npts), q = (int)(t % m.npts);
  int r = 0;
  while (q >= m.ps_off[r + 1]) r++;
  int g = m.ps_gln[r], kt = q - m.ps_off[r], k1 = kt / g, k2 = kt % g;
  const GLTables& G = gl();
  double X[27], x[3], T1[3], T2[3], N[3], p[9];
  load_X(m, e, X);
  double x1 = G.x11[g][k1], x2 = G.x11[g][k2], w1 = G.w11[g][k1], w2 = G.w11[g][k2];
  q9_geom(X, x1, x2, x, T1, T2);
  cross3(T1, T2, N);
  double j = sqrt(dot3(N, N));
  phi_q9(x1, x2, p);
  double* o = ps + t * kPsStride;
  for (int c = 0; c < 3; c++) { o[c] = x[c]; o[3 + c] = N[c] / j; }
  for (int n = 0; n < 9; n++) o[6 + n] = p[n] * j * w1 * w2;
  o[15] = 0;
}

__global__ void k_node_normals(DevMesh m, double* nfn) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.nn) return;
  double s[3] = {0, 0, 0};
  for (int q = m.inc_ptr[i]; q < m.inc_ptr[i + 1]; q++) {
    double X[27], x[3], T1[3], T2[3], N[3];
    load_X(m, m.inc_e[q], X);
    int k = m.inc_k[q];
    q9_geom(X, kXiQ9[k][0], kXiQ9[k][1], x, T1, T2);
    cross3(T1, T2, N);
    double j = sqrt(dot3(N, N));
    for (int c = 0; c < 3; c++) s[c] += N[c] / j;
  }
  double j = sqrt(dot3(s, s));
  for (int c = 0; c < 3; c++) nfn[3 * i + c] = s[c] / j;
}

// Integration class of pair (i, e): the decision sequence of fbem_bem_staela3d_sbie_auto
// (bem_staela3d.f90:549-587), which the harmonic version repeats (bem_harela3d.f90:1498-1536).
__device__ int classify_pair(const DevMesh& m, const QsiParams& qsp, const double* X,
                             const double* el, const double* x_i) {
  double rr[3] = {el[2] - x_i[0], el[3] - x_i[1], el[4] - x_i[2]};
  double rmin = sqrt(dot3(rr, rr)) - el[1], barxi[2] = {0, 0}, d;
  if (rmin > 4.0 * el[1]) {
    d = rmin / el[0];
  } else {
    nearest_q9(X, el[0], x_i, barxi, &rmin, &d);
    if (d <= 1e-12) return m.n_ps + 1;
  }
  int gln_near = qs_estimate(qsp, 0, 2, 5, d, barxi);
  int gln = max((int)el[5], gln_near);
  int gmax = m.ps_gln[m.n_ps - 1];
  if (gln <= gmax && gln_near > 0) {
    for (int r = 0; r < m.n_ps; r++)
      if (m.ps_gln[r] >= gln) return r;
  }
  return m.n_ps;
}

// Features of a pair for the learned rule choice (MFB_QNET_INPUTS floats):
//   log10(d) / 2 with d as sbie_auto defines it (ball distance over cl for far pairs, nearest
//   point distance over cl otherwise), log2(aspect), curvature, k cl / 4, max(|barxi_1|, |barxi_2|)
__device__ void pair_features(const double* X, const double* el, const double* x_i, double kw,
                              float* f) {
  double rr[3] = {el[2] - x_i[0], el[3] - x_i[1], el[4] - x_i[2]};
  double rmin = sqrt(dot3(rr, rr)) - el[1], barxi[2] = {0, 0}, d;
  if (rmin > 4.0 * el[1]) d = rmin / el[0];
  else nearest_q9(X, el[0], x_i, barxi, &rmin, &d);
  f[0] = (float)(0.5 * log10(fmax(d, 1e-12)));
  f[1] = (float)log2(el[6]);
  f[2] = (float)el[7];
  f[3] = (float)(0.25 * kw * el[0]);
  f[4] = (float)fmax(fabs(barxi[0]), fabs(barxi[1]));
}

__global__ void k_pair_features(DevMesh m, const int* pairs, int np, double kw, float* feats) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const int pid = pairs[q], e = pid / m.ncp, cp = pid % m.ncp;
  double X[27], el[8];
  load_X(m, e, X);
  for (int c = 0; c < 8; c++) el[c] = m.elem[8 * e + c];
  pair_features(X, el, m.xc + 3 * cp, kw, feats + (long long)q * MFB_QNET_INPUTS);
}

// features of every pair (pair id = e * ncp + cp), k h left at 0 (set per assembly by k_set_kh)
__global__ void k_all_features(DevMesh m, float* feats) {
  const long long pid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pid >= (long long)m.ne * m.ncp) return;
  const int e = (int)(pid / m.ncp), cp = (int)(pid % m.ncp);
  double X[27], el[8];
  load_X(m, e, X);
  for (int c = 0; c < 8; c++) el[c] = m.elem[8 * e + c];
  pair_features(X, el, m.xc + 3 * cp, 0.0, feats + pid * MFB_QNET_INPUTS);
}

__global__ void k_set_kh(DevMesh m, float* feats, double kw) {
  const long long pid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pid >= (long long)m.ne * m.ncp) return;
  feats[pid * MFB_QNET_INPUTS + 3] = (float)(0.25 * kw * m.elem[8 * (pid / m.ncp)]);
}

// integration class of every pair from the network's class (singular pairs keep theirs): class c <
// MFB_QNET_CLASSES - 1 -> the smallest precalculated rule of order >= c + 2, moved on to promo[]
// of that rule (the judge's promotions; identity without); otherwise, or when no precalculated
// rule is that long, the adaptive rule.  counts: pairs per (element, class).
__global__ void __launch_bounds__(256) k_learned_map(DevMesh m, const unsigned char* cls_ref,
                                                     const int* qcls, const int* promo,
                                                     unsigned char* lcls, int* counts) {
  const int e = blockIdx.x;
  __shared__ int hist[kMaxPrecalsets + 2];
  if (threadIdx.x < kMaxPrecalsets + 2) hist[threadIdx.x] = 0;
  __syncthreads();
  for (int cp = threadIdx.x; cp < m.ncp; cp += blockDim.x) {
    const long long pid = (long long)e * m.ncp + cp;
    int c = cls_ref[pid];
    if (c != m.n_ps + 1) {
      const int q = qcls[pid];
      c = m.n_ps;
      if (q < MFB_QNET_CLASSES - 1)
        for (int r = 0; r < m.n_ps; r++)
          if (m.ps_gln[r] >= q + 2) { c = promo[r]; break; }
    }
    lcls[pid] = (unsigned char)c;
    atomicAdd(&hist[c], 1);
  }
  __syncthreads();
  if (threadIdx.x < m.n_ps + 2) counts[(long long)e * (kMaxPrecalsets + 2) + threadIdx.x] = hist[threadIdx.x];
}

// pairs at positions pos[k] of a list
__global__ void k_gather_list(const int* lists, const long long* pos, int k, int* out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < k) out[q] = lists[pos[q]];
}

__global__ void __launch_bounds__(256) k_classify(DevMesh m, const QsiParams* __restrict__ qsp_p,
                                                  unsigned char* cls, int* counts) {
  const QsiParams& qsp = *qsp_p;
  const int e = blockIdx.x;
  __shared__ int hist[kMaxPrecalsets + 2];
  if (threadIdx.x < kMaxPrecalsets + 2) hist[threadIdx.x] = 0;
  __syncthreads();
  double X[27], el[8];
  load_X(m, e, X);
  for (int c = 0; c < 8; c++) el[c] = m.elem[8 * e + c];
  for (int cp = threadIdx.x; cp < m.ncp; cp += blockDim.x) {
    int c = classify_pair(m, qsp, X, el, m.xc + 3 * cp);
    cls[(long long)e * m.ncp + cp] = (unsigned char)c;
    atomicAdd(&hist[c], 1);
  }
  __syncthreads();
  if (threadIdx.x < m.n_ps + 2) counts[(long long)e * (kMaxPrecalsets + 2) + threadIdx.x] = hist[threadIdx.x];
}

// pair ids e*ncp+cp into per-class lists; offsets[e*(kMaxPrecalsets+2)+c] is where element e's
// pairs of class c start.  Inside an (element, class) run the points keep their order (wave ballots
// and per-wave prefix counts, no atomics), so a run of consecutive points has consecutive rows.
__global__ void __launch_bounds__(256) k_fill_lists(DevMesh m, const unsigned char* cls,
                                                    const long long* offsets, int* lists) {
  constexpr int S = kMaxPrecalsets + 2, W = 256 / kWave;
  const int e = blockIdx.x, lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  __shared__ int cursor[S];
  __shared__ int wcnt[W][S];
  if (threadIdx.x < S) cursor[threadIdx.x] = 0;
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int c0 = 0; c0 < m.ncp; c0 += 256) {
    const int cp = c0 + threadIdx.x;
    const int c = cp < m.ncp ? cls[(long long)e * m.ncp + cp] : -1;
    int mine = 0;
    for (int q = 0; q < m.n_ps + 2; q++) {
      const unsigned long long mask = __ballot(c == q);
      if (c == q) mine = __popcll(mask & below);
      if (lane == 0) wcnt[wave][q] = __popcll(mask);
    }
    __syncthreads();
    if (c >= 0) {
      int pos = cursor[c] + mine;
      for (int w = 0; w < wave; w++) pos += wcnt[w][c];
      lists[offsets[(long long)e * S + c] + pos] = e * m.ncp + cp;
    }
    __syncthreads();
    if (threadIdx.x < m.n_ps + 2) {
      int t = 0;
      for (int w = 0; w < W; w++) t += wcnt[w][threadIdx.x];
      cursor[threadIdx.x] += t;
    }
    __syncthreads();
  }
}

// Tiles of the scatter kernel (k_regular_tile): element e's run of precalculated rule c is cut
// into tiles of at most kTileP consecutive list entries; toff[e*(kMaxPrecalsets+2)+c] is the
// run's first tile.  A tile is (first list entry, count | rule << 8).
constexpr int kTileP = 14;
__global__ void k_fill_tiles(DevMesh m, const int* counts, const long long* offsets,
                             const long long* toff, int2* tiles) {
  constexpr int S = kMaxPrecalsets + 2;
  const int e = blockIdx.x;
  for (int c = 0; c < m.n_ps; c++) {
    const long long k = (long long)e * S + c;
    const int cnt = counts[k];
    const long long o = offsets[k], t0 = toff[k];
    for (int j = threadIdx.x; j * kTileP < cnt; j += blockDim.x)
      tiles[t0 + j] = make_int2((int)(o + (long long)j * kTileP),
                                min(kTileP, cnt - j * kTileP) | (c << 8));
  }
}

// classes of listed pairs only (a batch of independent pairs)
__global__ void k_classify_listed(DevMesh m, const QsiParams* __restrict__ qsp_p, const int* pids,
                                  int np, unsigned char* cls) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const int pid = pids[q], e = pid / m.ncp, cp = pid % m.ncp;
  double X[27], el[8];
  load_X(m, e, X);
  for (int c = 0; c < 8; c++) el[c] = m.elem[8 * e + c];
  cls[q] = (unsigned char)classify_pair(m, *qsp_p, X, el, m.xc + 3 * cp);
}

#pragma clang fp contract(fast)

// ------------------------------------------------------------------------------------------------
// assembly kernels
// ------------------------------------------------------------------------------------------------
__global__ void k_reset(long long* counters, int* err) {
  if (threadIdx.x < 2) counters[threadIdx.x] = 0;
  if (threadIdx.x == 0) *err = 0;
}

template <typename T>
__global__ void k_zero(T* A, long long lda, int n, T* b) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long total = (long long)n * n;
  if (lda == n) {   // contiguous: a flat fill, no index division
    for (; t < total; t += (long long)gridDim.x * blockDim.x) A[t] = zero_of(T());
  } else {
    for (; t < total; t += (long long)gridDim.x * blockDim.x) {
      long long col = t / n, row = t % n;
      A[row + lda * col] = zero_of(T());
    }
  }
  t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = zero_of(T());
}

template <typename T>
__global__ void k_fill_zero(T* a, long long n) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) a[q] = zero_of(T());
}

// LDS of one k_regular workgroup with `items` (pair, point) items: 18 kernel values and 9 weights
template <bool H>
inline size_t kreg_lds_bytes(int items) {
  return (size_t)items * (18 * sizeof(typename Val<H>::T) + 9 * sizeof(double));
}

// Precalculated-rule pairs (ext_pre: bem_staela3d.f90:592-650, bem_harela3d.f90:628-700) of one
// rule, 64/P pairs per wave: lanes evaluate the kernel at the (pair, point) items into LDS, then each
// lane sums one (pair, kind, c = 3 l + k) row over the rule's points for three element nodes.
// MODE kBlocks: write each pair's 9x3x3 h and g blocks to blocks[162 (blk_base + position) ...];
// MODE kVector: add the pair's contribution to the interior displacement of its point (vec).
constexpr int kBlocks = 1, kVector = 2;
template <bool H, int MODE>
__global__ void __launch_bounds__(kWave) k_regular(DevMesh m, const KernelParams* __restrict__ kpp,
                                                   const int* list, int count, int rule,
                                                   typename Val<H>::T* blocks, long long blk_base,
                                                   VecOut<typename Val<H>::T> vec) {
  typedef typename Val<H>::T T;
  constexpr bool VEC = MODE == kVector;
  __shared__ T vred[kWave * 3];  // MODE kVector: per (pair of the batch, l) sums
  const KernelParams& kp = *kpp;
  const int g = m.ps_gln[rule], P = g * g, B = P >= kWave ? 1 : kWave / P;
  const int base = blockIdx.x * B;
  const int nb = min(B, count - base);
  const int lane = threadIdx.x;
  // LDS sized by the launch for this rule's nb * P items: (pair, point) x (ft[9], fu[9]), then
  // (pair, point) x phi_n J w[9] -- a small rule's workgroup holds less and more of them share a CU
  extern __shared__ double kreg_lds[];
  T* sfs = reinterpret_cast<T*>(kreg_lds);
  double* sw = kreg_lds + (sizeof(T) / sizeof(double)) * 18 * (B * P);
  const int off = m.ps_off[rule];
  for (int it = lane; it < nb * P; it += kWave) {
    int bb = it / P, p = it % P;
    int pid = list[base + bb];
    int e = pid / m.ncp, cp = pid % m.ncp;
    const double* q = m.ps + ((long long)e * m.npts + off + p) * kPsStride;
    double x[3] = {q[0], q[1], q[2]}, n[3] = {q[3], q[4], q[5]};
    double xi[3] = {m.xc[3 * cp], m.xc[3 * cp + 1], m.xc[3 * cp + 2]};
    ext_kernel_row(kp, x, n, xi, &sfs[it * 18]);
    for (int c = 0; c < 9; c++) sw[it * 9 + c] = q[6 + c];
  }
  if (VEC)
    for (int q = lane; q < 3 * nb; q += kWave) vred[q] = zero_of(T());
  __syncthreads();
  // one lane per (pair, kind, c = 3 l + k, three element nodes): sums over the rule's points, each
  // kernel value read once and used for the three nodes
  for (int it = lane; it < nb * 54; it += kWave) {
    const int bb = it / 54, o = it % 54, kind = o / 27, c = (o % 27) / 3, n0 = 3 * (o % 3);
    const int pid = list[base + bb];
    const int e = pid / m.ncp, cp = pid % m.ncp;
    const T* f = &sfs[bb * P * 18 + kind * 9 + c];
    const double* wq = &sw[bb * P * 9 + n0];
    T acc[3] = {zero_of(T()), zero_of(T()), zero_of(T())};
    for (int p = 0; p < P; p++) {
      const T fv = f[p * 18];
      acc[0] += fv * wq[p * 9];
      acc[1] += fv * wq[p * 9 + 1];
      acc[2] += fv * wq[p * 9 + 2];
    }
    const int il = c / 3, ik = c % 3;
    if (VEC) {
      T sum = zero_of(T());
#pragma unroll
      for (int q = 0; q < 3; q++) {
        const int sn = m.conn[9 * e + n0 + q];
        if (kind == 0) sum = sum - scale_h(kp, acc[q]) * vec.u[3 * sn + ik];
        else sum = sum + scale_g(kp, acc[q]) * vec.t[3 * sn + ik];
      }
      lds_add(&vred[3 * bb + il], sum);
      continue;
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const T v = kind ? scale_g(kp, acc[q]) : scale_h(kp, acc[q]);
      blocks[162LL * (blk_base + base + bb) + 81 * kind + n0 + q + 9 * il + 27 * ik] = v;
    }
  }
  if (VEC) {
    __syncthreads();
    for (int q = lane; q < 3 * nb; q += kWave) {
      const int cp = list[base + q / 3] % m.ncp;
      atomic_add(&vec.out[3 * cp + q % 3], vred[q]);
    }
  }
}

// Precalculated-rule pairs scattered into A and b (ext_pre: bem_staela3d.f90:592-650,
// bem_harela3d.f90:628-700; scatter: src/assemble_bem_staela_equation.f90,
// src/assemble_bem_harela_equation.f90), one tile per workgroup: up to kTileP pairs of one element
// and one rule, consecutive entries of the class list (ascending collocation points, so mostly
// consecutive matrix rows).  The rule's P points run in chunks of PC <= kTileG points
// (ceil(P / kTileG) chunks of near-equal size):
//   kernel:   thread (slot, g) evaluates the 18 kernel values of pair `slot` at point p0 + g into
//             LDS, F[value][g][slot];
//   contract: thread (kind, ik, slot, il) -- one matrix row 3 crow + il of its pair, kind 0 t* -> h,
//             1 u* -> g, direction ik -- sums its kernel value times the weights phi_n J w of the
//             9 element nodes over the points, in rule order (weights: uniform loads).
// Each thread then adds its 9 entries to A (or b).  Consecutive threads hold consecutive rows of
// one column.  The tiles of one launch belong to elements of one colour (no shared node): two of
// them touch the same entry only if their points share rows (MCA points of one node), which add
// atomically.  Chunks of at most 13 points keep F within 53 KB (harmonic), so three workgroups
// share a CU's LDS; registers are held to 168 per thread to match (three waves per SIMD: ~11 %
// faster than two with 18-point chunks; a 25-point rule still runs in two chunks).
constexpr int kTileG = 13, kTileThreads = 256, kTileRows = 18 * kTileP;
static_assert(kTileRows <= kTileThreads && kTileP * kTileG <= kTileThreads, "tile shape");
template <bool H>
__global__ void __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(3, 3)))
    k_regular_tile(DevMesh m, const KernelParams* __restrict__ kpp, const int* __restrict__ list,
                   const int2* __restrict__ tiles, typename Val<H>::T* A, long long lda,
                   typename Val<H>::T* b) {
  typedef typename Val<H>::T T;
  constexpr int kStride = kTileG * kTileP + 1;  // values of one kind/component (padded)
  __shared__ T F[18 * kStride];
  const KernelParams& kp = *kpp;
  const int2 tile = tiles[blockIdx.x];
  const int cnt = tile.y & 0xff, rule = tile.y >> 8;
  const int t = threadIdx.x;
  const int e = list[tile.x] / m.ncp;
  const int P = m.ps_gln[rule] * m.ps_gln[rule];
  const int nch = (P + kTileG - 1) / kTileG, PC = (P + nch - 1) / nch;
  const double* pe = m.ps + ((long long)e * m.npts + m.ps_off[rule]) * kPsStride;
  // kernel evaluation item of this thread
  const int aslot = t % kTileP, ag = t / kTileP;
  const bool aact = aslot < cnt && ag < PC;
  double xi[3] = {0.0, 0.0, 0.0};
  if (aact) {
    const int cp = list[tile.x + aslot] % m.ncp;
    xi[0] = m.xc[3 * cp];
    xi[1] = m.xc[3 * cp + 1];
    xi[2] = m.xc[3 * cp + 2];
  }
  // contraction row of this thread: t = ((kind * 3 + ik) * kTileP + slot) * 3 + il
  const int il = t % 3, slot = (t / 3) % kTileP, kk = min(t / (3 * kTileP), 5);
  const int kind = kk / 3, ik = kk % 3;
  const bool live = t < kTileRows && slot < cnt;  // threads past kTileRows contract row 5 idly
  const T* fr = &F[(kind * 9 + 3 * il + ik) * kStride + slot];
  T acc[9];
#pragma unroll
  for (int n = 0; n < 9; n++) acc[n] = zero_of(T());
  for (int p0 = 0; p0 < P; p0 += PC) {
    const int np = min(PC, P - p0);
    __syncthreads();  // the previous chunk is consumed
    if (aact && ag < np) {
      const double* q = pe + (long long)(p0 + ag) * kPsStride;
      const double x[3] = {q[0], q[1], q[2]}, nv[3] = {q[3], q[4], q[5]};
      ext_kernel_row(kp, x, nv, xi, &F[ag * kTileP + aslot], kStride);
    }
    __syncthreads();
    for (int g = 0; g < np; g++) {
      const double* w = pe + (long long)(p0 + g) * kPsStride + 6;
      const T f = fr[g * kTileP];
#pragma unroll
      for (int n = 0; n < 9; n++) acc[n] += f * w[n];
    }
  }
  if (!live) return;
  const int cp = list[tile.x + slot] % m.ncp;
  const long long row = 3LL * m.crow[cp] + il;
  const bool shared = m.cshared[cp] != 0;
  T bs = zero_of(T());
  bool has_b = false;
#pragma unroll
  for (int n = 0; n < 9; n++) {
    const T v = kind ? scale_g(kp, acc[n]) : scale_h(kp, acc[n]);
    const long long col = 3LL * m.conn[9 * e + n] + ik;
    const int ct = m.ctype[col];
    const double cv = m.cval[col];
    if (kind == 0) {
      if (ct != 0) {
        if (shared) atomic_add(&A[row + lda * col], v);
        else A[row + lda * col] += v;
      } else if (cv != 0.0) {
        bs = bs - v * cv;
        has_b = true;
      }
    } else {
      if (ct == 0) {
        if (shared) atomic_add(&A[row + lda * col], -v);
        else A[row + lda * col] += -v;
      } else if (cv != 0.0) {
        bs = bs + (ct == 1 ? v * cv : v * (cv * m.nfn[col]));
        has_b = true;
      }
    }
  }
  if (has_b) atomic_add(&b[row], bs);
}

// Interior-displacement contribution of one pair held as the per-lane outputs o = lane + 64 t
// (o = kind 81 + c 9 + n, before the kernel constants) of the adaptive and singular kernels.
template <typename T>
__device__ void vec_reduce(const DevMesh& m, const KernelParams& kp, int e, int cp, const T* acc,
                           const VecOut<T>& vec, int lane) {
  __shared__ T red[3];
  if (lane < 3) red[lane] = zero_of(T());
  __syncthreads();
  T part[3] = {zero_of(T()), zero_of(T()), zero_of(T())};
  for (int t = 0; t < 3; t++) {
    const int o = lane + kWave * t;
    if (o >= 162) continue;
    const int kind = o / 81, c = (o % 81) / 9, n = o % 9, il = c / 3, ik = c % 3;
    const int sn = m.conn[9 * e + n];
    if (kind == 0) part[il] = part[il] - scale_h(kp, acc[t]) * vec.u[3 * sn + ik];
    else part[il] = part[il] + scale_g(kp, acc[t]) * vec.t[3 * sn + ik];
  }
  for (int il = 0; il < 3; il++) lds_add(&red[il], part[il]);
  __syncthreads();
  if (lane < 3) atomic_add(&vec.out[3 * cp + lane], red[lane]);
}

// Adaptive quasi-singular pairs: fbem_bem_staela3d_sbie_ext_adp (bem_staela3d.f90:973-1096,
// bem_harela3d.f90:1050-1172) with ext_st leaves (:654-969, :702-1048).
struct Leaf {
  double xs[8];
  double barxi[2];
  double barr;
  int gln;
  int pad;
};

template <bool H>
__global__ void __launch_bounds__(kWave) k_adaptive(DevMesh m, const QsiParams* __restrict__ qsp_p,
                                                    const KernelParams* __restrict__ kpp,
                                                    const int* list, int count,
                                                    typename Val<H>::T* A, long long lda,
                                                    typename Val<H>::T* b, long long* counters,
                                                    int* err, typename Val<H>::T* blocks,
                                                    long long blk_base,
                                                    VecOut<typename Val<H>::T> vec) {
  typedef typename Val<H>::T T;
  const QsiParams& qsp = *qsp_p;
  const KernelParams& kp = *kpp;
  const int lane = threadIdx.x;
  const int pid = list[blockIdx.x];
  const int e = pid / m.ncp, cp = pid % m.ncp;
  __shared__ Leaf leaves[kMaxLeaves];
  __shared__ double stk[kStack][9];  // xs(8) + level
  __shared__ int nleaves, sp_sh;
  __shared__ bool ovf;
  __shared__ T sfs[kChunk * 18];
  __shared__ double sw[kChunk * 9];
  double X[27], el[8];
  load_X(m, e, X);
  for (int c = 0; c < 8; c++) el[c] = m.elem[8 * e + c];
  const double xi_c[3] = {m.xc[3 * cp], m.xc[3 * cp + 1], m.xc[3 * cp + 2]};
  const int gln_far = (int)el[5];
  if (lane == 0) {
    const double root[8] = {-1, -1, 1, -1, 1, 1, -1, 1};
    for (int c = 0; c < 8; c++) stk[0][c] = root[c];
    stk[0][8] = 1;
    sp_sh = 1;
    ovf = false;
  }
  const GLTables& G = gl();
  // lane owns outputs o = lane + 64 t, o = kind*81 + c*9 + n
  T acc[3] = {zero_of(T()), zero_of(T()), zero_of(T())};
  long long npts = 0;
  int nl_total = 0;
  // The walk (lane 0) stops after kMaxLeaves leaves, all lanes integrate them, and the walk goes on
  // from its stack: leaves are integrated in the recursion's order and their number is unbounded.
  bool more = true;
  while (more) {
    __syncthreads();
    if (lane == 0) {
      int sp = sp_sh, nl = 0;
      bool overflow = false;
      while (sp > 0 && nl < kMaxLeaves) {
        sp--;
        double xs[8];
        for (int c = 0; c < 8; c++) xs[c] = stk[sp][c];
        int ks = (int)stk[sp][8];
        double barxi[2], rmin, d;
        if (ks == 1) {
          nearest_q9(X, el[0], xi_c, barxi, &rmin, &d);
        } else {
          double Xs[27];
          subdiv_q9(X, xs, Xs);
          double cl = clen_q9(Xs, 1e-12);
          nearest_q9(Xs, cl, xi_c, barxi, &rmin, &d);
        }
        int gln_near = qs_estimate(qsp, 1, 2, 5, d, barxi);
        bool sub = false;
        if (ks == m.ns_max) {
          if (gln_near == 0) gln_near = 30;
        } else if (gln_near == 0) {
          sub = true;
        }
        if (sub) {
          // the stack holds at most 3 ns_max + 1 entries (ns_max is bounded at mesh creation)
          if (sp + 4 > kStack) { overflow = true; break; }
          double c0 = 0.25 * (xs[0] + xs[2] + xs[4] + xs[6]), c1 = 0.25 * (xs[1] + xs[3] + xs[5] + xs[7]);
          double m01[2] = {0.5 * (xs[0] + xs[2]), 0.5 * (xs[1] + xs[3])};
          double m12[2] = {0.5 * (xs[2] + xs[4]), 0.5 * (xs[3] + xs[5])};
          double m23[2] = {0.5 * (xs[4] + xs[6]), 0.5 * (xs[5] + xs[7])};
          double m03[2] = {0.5 * (xs[0] + xs[6]), 0.5 * (xs[1] + xs[7])};
          const double s1[8] = {xs[0], xs[1], m01[0], m01[1], c0, c1, m03[0], m03[1]};
          const double s2[8] = {m01[0], m01[1], xs[2], xs[3], m12[0], m12[1], c0, c1};
          const double s3[8] = {c0, c1, m12[0], m12[1], xs[4], xs[5], m23[0], m23[1]};
          const double s4[8] = {m03[0], m03[1], c0, c1, m23[0], m23[1], xs[6], xs[7]};
          // push in reverse so that sub-quad 1 is integrated first, as the recursion does
          const double* sq[4] = {s4, s3, s2, s1};
          for (int s = 0; s < 4; s++) {
            for (int c = 0; c < 8; c++) stk[sp][c] = sq[s][c];
            stk[sp][8] = ks + 1;
            sp++;
          }
        } else {
          Leaf& L = leaves[nl++];
          for (int c = 0; c < 8; c++) L.xs[c] = xs[c];
          L.barxi[0] = barxi[0];
          L.barxi[1] = barxi[1];
          L.barr = telles_barr(d);
          L.gln = max(gln_near, gln_far);
        }
      }
      if (overflow) {
        atomicExch(err, MFB_ERR_LIMIT);
        ovf = true;
        sp = 0;
        nl = 0;
      }
      sp_sh = sp;
      nleaves = nl;
    }
    __syncthreads();
    const int nl = nleaves;
    more = sp_sh > 0;
    nl_total += nl;
    for (int l = 0; l < nl; l++) {
      const Leaf L = leaves[l];
      const int gln = L.gln, P = gln * gln;
      double tp1[4], tp2[4];
      telles11(L.barxi[0], L.barr, tp1);
      telles11(L.barxi[1], L.barr, tp2);
      npts += P;
      for (int start = 0; start < P; start += kChunk) {
        int p = start + lane;
        if (lane < kChunk && p < P) {
          int k1 = p / gln, k2 = p % gln;
          double xp0, xp1, jt1, jt2;
          telles_xj(tp1, G.x11[gln][k1], &xp0, &jt1);
          telles_xj(tp2, G.x11[gln][k2], &xp1, &jt2);
          double w1 = G.w11[gln][k1], w2 = G.w11[gln][k2];
          // xip -> xi through the quad4 map of the sub-element
          double gp[4] = {0.25 * (1.0 - xp0) * (1.0 - xp1), 0.25 * (1.0 + xp0) * (1.0 - xp1),
                          0.25 * (1.0 + xp0) * (1.0 + xp1), 0.25 * (1.0 - xp0) * (1.0 + xp1)};
          double b2 = 0.25 * xp1, b3 = 0.25, c2 = 0.25 * xp0;
          double d1[4] = {b2 - b3, -(b2 - b3), b2 + b3, -(b2 + b3)};
          double d2[4] = {c2 - b3, -c2 - b3, -(-c2 - b3), -(c2 - b3)};
          double xi0 = 0, xi1 = 0, a0 = 0, a1 = 0, a2 = 0, a3 = 0;
          for (int kk = 0; kk < 4; kk++) {
            xi0 += gp[kk] * L.xs[2 * kk]; xi1 += gp[kk] * L.xs[2 * kk + 1];
            a0 += d1[kk] * L.xs[2 * kk]; a1 += d1[kk] * L.xs[2 * kk + 1];
            a2 += d2[kk] * L.xs[2 * kk]; a3 += d2[kk] * L.xs[2 * kk + 1];
          }
          double js = a0 * a3 - a1 * a2;
          double x[3], T1[3], T2[3], N[3], n[3], ph[9];
          q9_geom(X, xi0, xi1, x, T1, T2);
          cross3(T1, T2, N);
          double jg = sqrt(dot3(N, N));
          for (int c = 0; c < 3; c++) n[c] = N[c] / jg;
          double jw = jg * js * jt1 * jt2 * w1 * w2;
          phi_q9(xi0, xi1, ph);
          ext_kernel_row(kp, x, n, xi_c, &sfs[lane * 18]);
          for (int c = 0; c < 9; c++) sw[lane * 9 + c] = ph[c] * jw;
        }
        __syncthreads();
        int np = min(kChunk, P - start);
        for (int t = 0; t < 3; t++) {
          int o = lane + kWave * t;
          if (o < 162) {
            int kind = o / 81, c = (o % 81) / 9, n = o % 9;
            for (int pp = 0; pp < np; pp++) acc[t] += sfs[pp * 18 + kind * 9 + c] * sw[pp * 9 + n];
          }
        }
        __syncthreads();
      }
    }
  }
  const int nl = nl_total;
  if (vec.out) {
    vec_reduce(m, kp, e, cp, acc, vec, lane);
  } else
  // each lane scatters its outputs directly (single values)
  for (int t = 0; t < 3; t++) {
    int o = lane + kWave * t;
    if (o < 162) {
      int kind = o / 81, c = (o % 81) / 9, n = o % 9;
      T v = kind ? scale_g(kp, acc[t]) : scale_h(kp, acc[t]);
      int il = c / 3, ik = c % 3;
      if (blocks) {
        // a pair whose subdivision overflowed is marked (NaN) rather than given a wrong value
        blocks[162LL * (blk_base + blockIdx.x) + 81 * kind + n + 9 * il + 27 * ik] =
            ovf ? from_real<T>(__builtin_nan("")) : v;
        continue;
      }
      int sn = m.conn[9 * e + n];
      long long row = 3LL * m.crow[cp] + il, col = 3LL * sn + ik;
      int ct = m.ctype[3 * sn + ik];
      double cv = m.cval[3 * sn + ik];
      if (kind == 0) {
        if (ct == 0) atomic_add(&b[row], -(v * cv));
        else atomic_add(&A[row + lda * col], v);
      } else {
        if (ct == 0) atomic_add(&A[row + lda * col], -v);
        else if (ct == 1) atomic_add(&b[row], v * cv);
        else atomic_add(&b[row], v * (cv * m.nfn[3 * sn + ik]));
      }
    }
  }
  if (lane == 0) {
    atomicAdd((unsigned long long*)&counters[0], (unsigned long long)npts);
    atomicAdd((unsigned long long*)&counters[1], (unsigned long long)nl);
  }
}

// Line integral eps_ijk e_k . t / r over one line3 edge, adaptive Telles + bisection
// (fbem_bem_staela3d_sbie_int_li, bem_staela3d.f90:2245-2376).  Returns the 3 independent
// entries (1,2), (1,3), (2,3).
__device__ __attribute__((noinline)) void line_integral(const double* Xl, const double* x_i, const QsiParams& q, int ns,
                              double* out) {
  const GLTables& G = gl();
  double stk[40][3];
  int sp = 0;
  stk[0][0] = -1.0; stk[0][1] = 1.0; stk[0][2] = 1;
  sp = 1;
  out[0] = out[1] = out[2] = 0;
  while (sp > 0) {
    sp--;
    double xs[2] = {stk[sp][0], stk[sp][1]};
    int ks = (int)stk[sp][2];
    double Xs[9];
    if (ks == 1) {
      for (int c = 0; c < 9; c++) Xs[c] = Xl[c];
    } else {
      subdiv_l3(Xl, xs, Xs);
    }
    double cl = length_l3(Xs, 1e-12), bx, rmin, d;
    nearest_l3(Xs, cl, x_i, &bx, &rmin, &d);
    int gln_near = qs_estimate(q, 1, 1, 1, d, &bx);
    bool sub = false;
    if (ks == ns) {
      if (gln_near == 0) gln_near = 30;
    } else if (gln_near == 0) {
      sub = true;
    }
    if (sub && sp + 2 <= 40) {
      double mid = 0.5 * (xs[0] + xs[1]);
      stk[sp][0] = mid; stk[sp][1] = xs[1]; stk[sp][2] = ks + 1; sp++;
      stk[sp][0] = xs[0]; stk[sp][1] = mid; stk[sp][2] = ks + 1; sp++;
    } else {
      int gln = max(gln_near, 5);
      double tp[4], t12 = 0, t13 = 0, t23 = 0;
      telles11(bx, telles_barr(d), tp);
      for (int kip = 0; kip < gln; kip++) {
        double xip, jt;
        telles_xj(tp, G.x11[gln][kip], &xip, &jt);
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
        double jg = sqrt(dot3(T, T));
        double t[3] = {T[0] / jg, T[1] / jg, T[2] / jg};
        double rv[3] = {x[0] - x_i[0], x[1] - x_i[1], x[2] - x_i[2]};
        double dr1 = 1.0 / sqrt(dot3(rv, rv));
        double jw = jg * js * jt * G.w11[gln][kip];
        t12 = t12 - dr1 * t[2] * jw;
        t13 = t13 + dr1 * t[1] * jw;
        t23 = t23 - dr1 * t[0] * jw;
      }
      out[0] += t12;
      out[1] += t13;
      out[2] += t23;
    }
  }
}

// Singular pairs: fbem_bem_staela3d_sbie_int (bem_staela3d.f90:1100-1381) and
// fbem_bem_harela3d_sbie_int (bem_harela3d.f90:1174-1472).
template <bool H>
__global__ void __launch_bounds__(kWave) k_singular(DevMesh m, const QsiParams* __restrict__ qsp_li_p,
                                                    const KernelParams* __restrict__ kpp,
                                                    const int* list, int count,
                                                    typename Val<H>::T* A, long long lda,
                                                    typename Val<H>::T* b, long long* counters,
                                                    typename Val<H>::T* blocks, long long blk_base,
                                                    VecOut<typename Val<H>::T> vec) {
  typedef typename Val<H>::T T;
  const QsiParams& qsp_li = *qsp_li_p;
  const KernelParams& kp = *kpp;
  const int lane = threadIdx.x;
  const int pid = list[blockIdx.x];
  const int e = pid / m.ncp, cp = pid % m.ncp;
  __shared__ T sfs[kChunk * 27];     // ft, fu, fc
  __shared__ double sw[kChunk * 18];  // phi*jw, (phi - phi_i)*jw
  __shared__ double shli[4][3];
  double X[27], el[8];
  load_X(m, e, X);
  for (int c = 0; c < 8; c++) el[c] = m.elem[8 * e + c];
  const double xc[3] = {m.xc[3 * cp], m.xc[3 * cp + 1], m.xc[3 * cp + 2]};
  double xi_i[2], rmin, d;
  if (m.cel[cp] == e) {
    // an MCA point inside this element: its local coordinates are known exactly
    xi_i[0] = m.cxi[2 * cp];
    xi_i[1] = m.cxi[2 * cp + 1];
  } else {
    nearest_q9(X, el[0], xc, xi_i, &rmin, &d);
  }
  double x_i[3], pf_i[9];
  q9_x(X, xi_i[0], xi_i[1], x_i);
  phi_q9(xi_i[0], xi_i[1], pf_i);
  int st[8], ngt[8], base[9], nst = -1;
  double th[16], thp[16];
  for (int k = 0; k < 9; k++)
    if (xi_i[0] == kXiQ9[k][0] && xi_i[1] == kXiQ9[k][1]) {
      const PolarSet& ps = c_polar[k];
      nst = ps.n;
      for (int s = 0; s < nst; s++) {
        st[s] = ps.st[s];
        ngt[s] = ps.ngt[s];
        th[2 * s] = ps.th[2 * s];
        th[2 * s + 1] = ps.th[2 * s + 1];
        thp[2 * s] = ps.thp[2 * s];
        thp[2 * s + 1] = ps.thp[2 * s + 1];
      }
      break;
    }
  if (nst < 0) {
    nst = polar_setup_q(xi_i, st, th, thp);
    for (int s = 0; s < nst; s++) ngt[s] = 5 + (int)round(25.0 * (th[2 * s + 1] - th[2 * s]) / kPi_2);
  }
  base[0] = 0;
  for (int s = 0; s < nst; s++) base[s + 1] = base[s] + ngt[s] * 15;
  const int P = base[nst];
  const GLTables& G = gl();
  T acc[3] = {zero_of(T()), zero_of(T()), zero_of(T())};
  for (int start = 0; start < P; start += kChunk) {
    int p = start + lane;
    if (lane < kChunk && p < P) {
      int s = 0;
      while (p >= base[s + 1]) s++;
      int q = p - base[s], kt = q / 15, kr = q % 15, nt = ngt[s];
      double jtp = thp[2 * s + 1] - thp[2 * s];
      double tpv = jtp * G.x01[nt][kt] + thp[2 * s];
      double theta, rhoij;
      polar_angular_q(xi_i, st[s], tpv, &theta, &rhoij);
      double sn, cs;
      sincos(theta, &sn, &cs);
      double rho = rhoij * G.x01[15][kr];
      double xi0 = xi_i[0] + rho * cs, xi1 = xi_i[1] + rho * sn;
      double x[3], T1[3], T2[3], N[3], n[3], ph[9];
      q9_geom(X, xi0, xi1, x, T1, T2);
      cross3(T1, T2, N);
      double jg = sqrt(dot3(N, N));
      for (int c = 0; c < 3; c++) n[c] = N[c] / jg;
      double jw = jg * rho * jtp * G.w01[nt][kt] * G.w01[15][kr];
      phi_q9(xi0, xi1, ph);
      sing_kernel_row(kp, x, n, x_i, &sfs[lane * 27]);
      for (int c = 0; c < 9; c++) {
        sw[lane * 18 + c] = ph[c] * jw;
        sw[lane * 18 + 9 + c] = (ph[c] - pf_i[c]) * jw;
      }
    }
    __syncthreads();
    int np = min(kChunk, P - start);
    for (int t = 0; t < 3; t++) {
      int o = lane + kWave * t;
      if (o < 162) {
        int kind = o / 81, c = (o % 81) / 9, n = o % 9;
        if (kind == 0) {
          for (int pp = 0; pp < np; pp++)
            acc[t] += sfs[pp * 27 + c] * sw[pp * 18 + n] + sfs[pp * 27 + 18 + c] * sw[pp * 18 + 9 + n];
        } else {
          for (int pp = 0; pp < np; pp++) acc[t] += sfs[pp * 27 + 9 + c] * sw[pp * 18 + n];
        }
      }
    }
    __syncthreads();
  }
  // Stokes line integrals over the edges that bound an integrated sub-triangle
  if (lane < 4) {
    bool integrate = false;
    for (int s = 0; s < nst; s++)
      if (st[s] == 2 * (lane + 1) - 1 || st[s] == 2 * (lane + 1)) integrate = true;
    double o3[3] = {0, 0, 0};
    if (integrate) {
      double Xl[9];
      for (int j = 0; j < 3; j++)
        for (int c = 0; c < 3; c++) Xl[3 * j + c] = X[3 * kEdgeQ9[lane][j] + c];
      line_integral(Xl, x_i, qsp_li, 16, o3);
    }
    shli[lane][0] = o3[0];
    shli[lane][1] = o3[1];
    shli[lane][2] = o3[2];
  }
  __syncthreads();
  double h12 = 0, h13 = 0, h23 = 0;
  for (int ke = 0; ke < 4; ke++) { h12 += shli[ke][0]; h13 += shli[ke][1]; h23 += shli[ke][2]; }
  // hli(il,ik): antisymmetric with (1,2)=h12, (1,3)=h13, (2,3)=h23
  const double hli[3][3] = {{0, h12, h13}, {-h12, 0, h23}, {-h13, -h23, 0}};
  if (vec.out) {
    T fin[3];
    for (int t = 0; t < 3; t++) {
      const int o = lane + kWave * t;
      fin[t] = acc[t];
      if (o < 81) {
        const int c = o / 9, n = o % 9;
        const double f = pf_i[n] * hli[c / 3][c % 3];
        if constexpr (H) fin[t] += mkz(kp.h.T2[0][0], kp.h.T2[0][1]) * f;
        else fin[t] += kp.ctet2 * f;
      }
    }
    vec_reduce(m, kp, e, cp, fin, vec, lane);
    if (lane == 0) atomicAdd((unsigned long long*)&counters[0], (unsigned long long)P);
    return;
  }
  for (int t = 0; t < 3; t++) {
    int o = lane + kWave * t;
    if (o < 162) {
      int kind = o / 81, c = (o % 81) / 9, n = o % 9;
      int il = c / 3, ik = c % 3;
      T v = acc[t];
      if (kind == 0) {
        double f = pf_i[n] * hli[il][ik];
        if constexpr (H) v += mkz(kp.h.T2[0][0], kp.h.T2[0][1]) * f;
        else v += kp.ctet2 * f;
      }
      v = kind ? scale_g(kp, v) : scale_h(kp, v);
      if (blocks) {
        blocks[162LL * (blk_base + blockIdx.x) + 81 * kind + n + 9 * il + 27 * ik] = v;
        continue;
      }
      int sn = m.conn[9 * e + n];
      long long row = 3LL * m.crow[cp] + il, col = 3LL * sn + ik;
      int ct = m.ctype[3 * sn + ik];
      double cv = m.cval[3 * sn + ik];
      if (kind == 0) {
        if (ct == 0) atomic_add(&b[row], -(v * cv));
        else atomic_add(&A[row + lda * col], v);
      } else {
        if (ct == 0) atomic_add(&A[row + lda * col], -v);
        else if (ct == 1) atomic_add(&b[row], v * cv);
        else atomic_add(&b[row], v * (cv * m.nfn[3 * sn + ik]));
      }
    }
  }
  if (lane == 0) atomicAdd((unsigned long long*)&counters[0], (unsigned long long)P);
}

// err[q * nr + r] = max(max|h - h_ref| / max|h_ref|, max|g - g_ref| / max|g_ref|) of pair q
// (one wave per pair; blocks of 162 values, h then g)
template <typename T>
__global__ void __launch_bounds__(kWave) k_block_err(const T* ref, const T* blk, int nr, int r,
                                                     double* err) {
  const long long q = blockIdx.x;
  const int lane = threadIdx.x;
  double dh = 0, mh = 0, dg = 0, mg = 0;
  for (int o = lane; o < 162; o += kWave) {
    const double a = mag(ref[162 * q + o]), dd = mag(blk[162 * q + o] - ref[162 * q + o]);
    if (o < 81) { mh = fmax(mh, a); dh = fmax(dh, dd); }
    else { mg = fmax(mg, a); dg = fmax(dg, dd); }
  }
  for (int off = 32; off > 0; off >>= 1) {
    dh = fmax(dh, __shfl_xor(dh, off));
    mh = fmax(mh, __shfl_xor(mh, off));
    dg = fmax(dg, __shfl_xor(dg, off));
    mg = fmax(mg, __shfl_xor(mg, off));
  }
  if (lane == 0) {
    const double r0 = mag(ref[162 * q]);
    // a reference marked NaN (overflowed subdivision) gives a NaN error
    err[q * nr + r] = r0 != r0 ? r0 : fmax(mh > 0 ? dh / mh : dh, mg > 0 ? dg / mg : dg);
  }
}

// label of pair q: class of the cheapest precalculated rule with err <= tau (class = order - 2),
// MFB_QNET_CLASSES - 1 (adaptive) when none is; written to labels[idx[q]]
__global__ void k_rule_labels(const double* err, int np, int nr, DevMesh m, double tau,
                              const int* idx, int* labels) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  int lab = MFB_QNET_CLASSES - 1;
  const double e0 = err[(long long)q * nr];
  if (e0 != e0) {
    lab = -1;  // no reference (subdivision overflow): no label
  } else {
    for (int r = 0; r < nr; r++)
      if (err[(long long)q * nr + r] <= tau) {
        lab = m.ps_gln[r] - 2;
        break;
      }
  }
  labels[idx[q]] = lab;
}

__global__ void k_fill_int(int* a, int n, int v) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) a[q] = v;
}

#pragma clang fp contract(off)

// Mantic free term (bem_staela3d.f90:228-405) gathered as in
// src/build_lse_mechanics_bem_staela.f90:210-549; I/2 for nodes interior to their first element.
__device__ bool mantic(int m_, double (*ln)[3], double (*lt)[3], double tol, double nu,
                       double* C) {
  // ln, lt indexed 0..m_+1 (0 and m_+1 are the wrap-around copies)
  double lti[kMaxIncid][3], theta[kMaxIncid];
  int tc[kMaxIncid];
  for (int ki = 1; ki <= m_ - 1; ki++) {
    double e1[3] = {lt[ki][0], lt[ki][1], lt[ki][2]}, e3[3] = {ln[ki][0], ln[ki][1], ln[ki][2]}, e2[3];
    e2[0] = e3[1] * e1[2] - e3[2] * e1[1];
    e2[1] = e3[2] * e1[0] - e3[0] * e1[2];
    e2[2] = e3[0] * e1[1] - e3[1] * e1[0];
    for (int kj = 1; kj <= m_; kj++) {
      lti[kj - 1][0] = e1[0] * lt[kj][0] + e1[1] * lt[kj][1] + e1[2] * lt[kj][2];
      lti[kj - 1][1] = e2[0] * lt[kj][0] + e2[1] * lt[kj][1] + e2[2] * lt[kj][2];
      lti[kj - 1][2] = e3[0] * lt[kj][0] + e3[1] * lt[kj][1] + e3[2] * lt[kj][2];
    }
    int nt = 0;
    for (int kj = ki + 1; kj <= m_; kj++)
      if (fabs(lti[kj - 1][2]) <= tol) tc[nt++] = kj;
    if (nt == 0) return false;
    for (int j = 0; j < nt; j++) {
      double a = atan2(lti[tc[j] - 1][1], lti[tc[j] - 1][0]);
      if (a < 0.0) a += 2.0 * kPi;
      theta[tc[j] - 1] = a;
    }
    double mint = theta[tc[0] - 1];
    int mk = tc[0];
    for (int j = 1; j < nt; j++)
      if (theta[tc[j] - 1] < mint) { mint = theta[tc[j] - 1]; mk = tc[j]; }
    for (int c = 0; c < 3; c++) {
      double t = lt[ki + 1][c]; lt[ki + 1][c] = lt[mk][c]; lt[mk][c] = t;
      t = ln[ki + 1][c]; ln[ki + 1][c] = ln[mk][c]; ln[mk][c] = t;
    }
  }
  for (int c = 0; c < 3; c++) {
    ln[0][c] = ln[m_][c]; lt[0][c] = lt[m_][c];
    ln[m_ + 1][c] = ln[1][c]; lt[m_ + 1][c] = lt[1][c];
  }
  double sum_a = 0;
  for (int ki = 1; ki <= m_; ki++) {
    double nx[3];
    nx[0] = ln[ki - 1][1] * ln[ki][2] - ln[ki - 1][2] * ln[ki][1];
    nx[1] = ln[ki - 1][2] * ln[ki][0] - ln[ki - 1][0] * ln[ki][2];
    nx[2] = ln[ki - 1][0] * ln[ki][1] - ln[ki - 1][1] * ln[ki][0];
    double s = nx[0] * lt[ki][0] + nx[1] * lt[ki][1] + nx[2] * lt[ki][2];
    if (s < 0.0) s = -1.0;
    if (s > 0.0) s = 1.0;
    double ndn = ln[ki - 1][0] * ln[ki][0] + ln[ki - 1][1] * ln[ki][1] + ln[ki - 1][2] * ln[ki][2];
    if (ndn > 1.0) ndn = 1.0;
    sum_a = sum_a + s * acos(ndn);
  }
  double cp = 1.0 / (4.0 * kPi) * (2.0 * kPi + sum_a);
  double sb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int ki = 1; ki <= m_; ki++) {
    double rr[3], rx[3];
    for (int c = 0; c < 3; c++) rr[c] = lt[ki + 1][c] - lt[ki][c];
    rx[0] = rr[1] * ln[ki][2] - rr[2] * ln[ki][1];
    rx[1] = rr[2] * ln[ki][0] - rr[0] * ln[ki][2];
    rx[2] = rr[0] * ln[ki][1] - rr[1] * ln[ki][0];
    for (int c1 = 0; c1 < 3; c1++)
      for (int c2 = 0; c2 < 3; c2++) sb[c1 + 3 * c2] = sb[c1 + 3 * c2] + rx[c1] * ln[ki][c2];
  }
  for (int c1 = 0; c1 < 3; c1++)
    for (int c2 = 0; c2 < 3; c2++) {
      C[c1 + 3 * c2] = -1.0 / (8.0 * kPi * (1.0 - nu)) * sb[c1 + 3 * c2];
      if (c1 == c2) C[c1 + 3 * c2] += cp;
    }
  return true;
}

template <typename T>
__global__ void k_freeterm(DevMesh m, double nu, T* A, long long lda, T* b, int* err) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.nn) return;
  int q0 = m.inc_ptr[i], q1 = m.inc_ptr[i + 1];
  if (q1 == q0 || m.nmca[i]) return;
  double C[9];
  if (m.inc_k[q0] == 8) {
    for (int c = 0; c < 9; c++) C[c] = (c % 4 == 0) ? 0.5 : 0.0;
  } else {
    int mm = q1 - q0;
    if (mm > kMaxIncid - 2) { atomicExch(err, MFB_ERR_LIMIT); return; }
    double ln[kMaxIncid][3], lt[kMaxIncid][3];
    for (int q = q0; q < q1; q++) {
      double X[27], x[3], T1[3], T2[3], N[3];
      load_X(m, m.inc_e[q], X);
      int k = m.inc_k[q];
      q9_geom(X, kXiQ9[k][0], kXiQ9[k][1], x, T1, T2);
      cross3(T1, T2, N);
      double j = sqrt(dot3(N, N));
      int s = q - q0 + 1;
      for (int c = 0; c < 3; c++) ln[s][c] = N[c] / j;
      // boundary tangent tbp at the node (geometry.f90:657-870)
      const double* Tt = (k == 0 || k == 4 || k == 2 || k == 6) ? T1 : T2;
      double sg = (k == 2 || k == 6 || k == 3 || k == 7) ? -1.0 : 1.0;
      double nt = sqrt(Tt[0] * Tt[0] + Tt[1] * Tt[1] + Tt[2] * Tt[2]);
      for (int c = 0; c < 3; c++) lt[s][c] = sg * (Tt[c] / nt);
    }
    double tol = (m.tol < 1e-12 || m.tol > 1e-3) ? 1e-6 : m.tol;
    if (!mantic(mm, ln, lt, tol, nu, C)) { atomicExch(err, MFB_ERR_GEOMETRY); return; }
  }
  for (int il = 0; il < 3; il++) {
    long long row = 3LL * i + il;
    for (int ik = 0; ik < 3; ik++) {
      long long col = 3LL * i + ik;
      double cv = C[il + 3 * ik];
      if (m.ctype[3 * i + ik] == 0) atomic_add(&b[row], from_real<T>(-(cv * m.cval[3 * i + ik])));
      else atomic_add(&A[row + lda * col], from_real<T>(cv));
    }
  }
}

// Free term of an MCA point: (1/2) u at the point, u interpolated on its element, i.e. (1/2)
// phi_n(xi_i) I on the element's nodes (src/build_lse_mechanics_bem_staela.f90:444-457), scattered by
// the BE-ordinary rule.
template <typename T>
__global__ void k_freeterm_mca(DevMesh m, T* A, long long lda, T* b) {
  const int cp = blockIdx.x * blockDim.x + threadIdx.x;
  if (cp >= m.ncp) return;
  const int e = m.cel[cp];
  if (e < 0) return;
  double ph[9];
  phi_q9(m.cxi[2 * cp], m.cxi[2 * cp + 1], ph);
  const int i = m.crow[cp];
  for (int n = 0; n < 9; n++) {
    const int sn = m.conn[9 * e + n];
    const double h = 0.5 * ph[n];
    for (int il = 0; il < 3; il++) {
      const long long row = 3LL * i + il, col = 3LL * sn + il;
      if (m.ctype[3 * sn + il] == 0) atomic_add(&b[row], from_real<T>(-(h * m.cval[3 * sn + il])));
      else atomic_add(&A[row + lda * col], from_real<T>(h));
    }
  }
}

#pragma clang fp contract(fast)

}  // namespace mfb

// ================================================================================================
// host side
// ================================================================================================
namespace mfb {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

static int upload_polar_table() {
  static const double xin[9][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}, {0, -1},
                                   {1, 0},   {0, 1},  {-1, 0}, {0, 0}};
  PolarSet t[9];
  std::memset(t, 0, sizeof t);
  for (int k = 0; k < 9; k++) {
    t[k].n = polar_setup_q(xin[k], t[k].st, t[k].th, t[k].thp);
    for (int s = 0; s < t[k].n; s++)
      t[k].ngt[s] = 5 + (int)std::round(25.0 * (t[k].th[2 * s + 1] - t[k].th[2 * s]) / kPi_2);
  }
  MFB_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_polar), t, sizeof t));
  return MFB_OK;
}

static int upload_gl_tables() {
  static bool done = false;
  static int device = -1;
  int dev = 0;
  MFB_HIP(hipGetDevice(&dev));
  if (done && dev == device) return MFB_OK;
  static GLTables t;
  host_gl_tables(&t.x11[0][0], &t.w11[0][0], &t.x01[0][0], &t.w01[0][0]);
  MFB_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_gl), &t, sizeof t));
  done = true;
  device = dev;
  return MFB_OK;
}

}  // namespace mfb

namespace mfb {
// scatter tiles of the precalculated classes (k_regular_tile), launched per element colour
struct TileSet {
  int2* d = nullptr;
  size_t cap = 0;                  // tiles the device buffer holds
  std::vector<long long> off, cnt;  // per colour: first tile, tiles
};
}  // namespace mfb

struct mfb_mesh_s {
  mfb::DevMesh dm;
  mfb::QsiParams qsp, qsp_li;
  int n_cls;
  std::vector<long long> cls_off;  // start of each class list in `lists`
  std::vector<long long> cls_cnt;
  // element colours: elements of one colour share no node, so the pairs of one colour write
  // disjoint entries of A and need no atomics; inside a class list the elements run in
  // (colour, element) order
  int n_colors = 0;
  std::vector<int> ecol, eorder;   // colour of each element; elements in (colour, index) order
  mfb::TileSet tiles;              // of `lists`
  // learned rule choice (mfb_mesh_set_qnet): network, static pair features, the classes and lists
  // of the last assembly, and the judge's result
  mfb_qnet_s* qnet = nullptr;
  float qconf = 0.99f;
  int judge_samples = 0;
  mfb::QsiParams* d_qsp_judge = nullptr;
  // the judge's action (mfb_mesh_set_judge_action): a precalculated rule whose judged pairs miss
  // the tolerance is replaced by the cheapest longer rule that meets it on them (or the adaptive
  // rule) for this and every later assembly of the mesh; promo[r] is that replacement
  int judge_promote = 0;
  std::vector<int> promo;
  int* d_promo = nullptr;
  long long judge_promoted = 0;     // rules promoted by the last assembly
  double judge_first_max = 0;       // the judge's largest error before those promotions
  bool in_redo = false;
  float* feats = nullptr;          // (ne * nn) x MFB_QNET_INPUTS
  int* qcls = nullptr;             // network class per pair
  unsigned char* lcls = nullptr;   // integration class per pair chosen with the network
  int* llists = nullptr;
  std::vector<long long> lcls_off, lcls_cnt;
  mfb::TileSet ltiles;             // of `llists`
  long long l_regular_points = 0;
  bool learned_last = false;
  long long judge_n = 0;
  double judge_max = 0, judge_mean = 0;
  long long judge_ref_n = 0;       // judge pairs where the reference estimator chose a
  double judge_ref_max = 0;        // precalculated rule, and that rule's largest error
  long long* judge_counters = nullptr;  // scratch of the judge's adaptive kernel
  int* judge_err = nullptr;
  // grow-only device workspaces of the per-assembly steps (lists, learned classes, judge): a
  // repeated assembly allocates and frees nothing
  void* ws[8] = {};
  size_t ws_cap[8] = {};
  double *x = nullptr, *cval = nullptr, *elem = nullptr, *ps = nullptr, *nfn = nullptr;
  int *conn = nullptr, *ctype = nullptr, *inc_ptr = nullptr, *inc_e = nullptr, *inc_k = nullptr;
  unsigned char* cls = nullptr;
  int* lists = nullptr;
  // collocation points (device) and host copies of their rows / elements / coordinates
  double *xc = nullptr, *cxi = nullptr;
  int *crow = nullptr, *cel = nullptr;
  unsigned char *cshared = nullptr, *nmca = nullptr;
  std::vector<int> h_crow, h_cel;
  std::vector<double> h_xc;
  int n_mca = 0;
  // the geometry and settings it was created from (interior point sets are built from them)
  std::vector<double> h_x, h_cval;
  std::vector<int> h_conn, h_ctype, h_ps;
  double h_tol = 1e-6;
  int* err = nullptr;
  long long* counters = nullptr;  // points, leaves of adaptive + singular kernels
  mfb::QsiParams* d_qsp = nullptr;     // estimator tables and kernel constants live in device memory:
  mfb::QsiParams* d_qsp_li = nullptr;  // they are too large to travel as kernel arguments
  mfb::KernelParams* d_kp = nullptr;
  long long decisions = 0;
  long long regular_points = 0;
  // HIP events around the kernel groups of the last assembly (on the caller's stream)
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  double last_ms[4] = {0, 0, 0, 0};  // regular, adaptive, singular, free term + zeroing
  int last_regular_launches = 0;
};

using namespace mfb;

extern "C" const char* mfb_last_error(void) { return g_err; }

template <typename Tp>
static int dalloc(Tp** p, size_t n) {
  MFB_HIP(hipMalloc((void**)p, sizeof(Tp) * (n ? n : 1)));
  return MFB_OK;
}

static void mesh_release(mfb_mesh_s* m) {
  for (hipEvent_t& e : m->ev)
    if (e) (void)hipEventDestroy(e);
  void* ps[] = {m->x,     m->cval,  m->elem,     m->ps,   m->nfn,      m->conn,
                m->ctype, m->inc_ptr, m->inc_e, m->inc_k, m->cls,  m->lists, m->err,
                m->counters, m->d_qsp, m->d_qsp_li, m->d_kp, m->d_qsp_judge, m->feats,
                m->qcls, m->lcls, m->llists, m->judge_counters, m->judge_err,
                m->xc, m->cxi, m->crow, m->cel, m->cshared, m->nmca,
                m->ws[0], m->ws[1], m->ws[2], m->ws[3], m->ws[4], m->ws[5], m->ws[6], m->ws[7],
                m->tiles.d, m->ltiles.d, m->d_promo};
  for (void* p : ps)
    if (p) (void)hipFree(p);
}

enum MeshWs { MW_OFF, MW_COUNTS, MW_REF, MW_BLK, MW_POS, MW_SAMPLE, MW_ERR, MW_PC };

template <typename P>
static int mesh_ws(mfb_mesh_s* m, int slot, size_t bytes, P** out) {
  if (m->ws_cap[slot] < bytes) {
    if (m->ws[slot]) (void)hipFree(m->ws[slot]);
    m->ws[slot] = nullptr;
    m->ws_cap[slot] = 0;
    MFB_HIP(hipMalloc(&m->ws[slot], bytes));
    m->ws_cap[slot] = bytes;
  }
  *out = reinterpret_cast<P*>(m->ws[slot]);
  return MFB_OK;
}

// Class lists of the pairs: class-major; inside a class the elements run in (colour, index) order.
// counts: device (ne x (kMaxPrecalsets + 2)) pairs of each class per element; lists: device
// (ne * nn); outputs the host offsets and counts per class and per (class, colour).
static int build_lists(mfb_mesh_t m, const unsigned char* cls, const int* counts, int* lists,
                       std::vector<long long>& cls_off, std::vector<long long>& cls_cnt,
                       TileSet& ts, long long& regular_points, hipStream_t s) {
  const DevMesh& d = m->dm;
  const int ne = d.ne, S = kMaxPrecalsets + 2, nc = m->n_colors;
  std::vector<int> hc((size_t)ne * S);
  MFB_HIP(hipMemcpyAsync(hc.data(), counts, sizeof(int) * ne * S, hipMemcpyDeviceToHost, s));
  MFB_HIP(hipStreamSynchronize(s));
  cls_off.assign(m->n_cls + 1, 0);
  cls_cnt.assign(m->n_cls, 0);
  for (int e = 0; e < ne; e++)
    for (int c = 0; c < m->n_cls; c++) cls_cnt[c] += hc[(size_t)e * S + c];
  for (int c = 0; c < m->n_cls; c++) cls_off[c + 1] = cls_off[c] + cls_cnt[c];
  // off: list start of each (element, class) run; toff: first scatter tile of each (element,
  // precalculated class) run, tiles ordered by colour, then by rule (longest first), then element
  std::vector<long long> off((size_t)2 * ne * S, 0);
  long long* toff = off.data() + (size_t)ne * S;
  std::vector<long long> run(cls_off.begin(), cls_off.end() - 1);
  for (int c = 0; c < m->n_cls; c++)
    for (int q = 0; q < ne; q++) {
      const int e = m->eorder[q];
      off[(size_t)e * S + c] = run[c];
      run[c] += hc[(size_t)e * S + c];
    }
  std::vector<int> rorder(d.n_ps);
  for (int r = 0; r < d.n_ps; r++) rorder[r] = r;
  std::stable_sort(rorder.begin(), rorder.end(),
                   [&](int a, int b) { return d.ps_gln[a] > d.ps_gln[b]; });
  ts.off.assign(nc, 0);
  ts.cnt.assign(nc, 0);
  long long nt = 0;
  for (int k = 0; k < nc; k++) {
    ts.off[k] = nt;
    for (int r : rorder)
      for (int q = 0; q < ne; q++) {
        const int e = m->eorder[q];
        if (m->ecol[e] != k) continue;
        toff[(size_t)e * S + r] = nt;
        nt += (hc[(size_t)e * S + r] + kTileP - 1) / kTileP;
      }
    ts.cnt[k] = nt - ts.off[k];
  }
  if ((size_t)nt > ts.cap) {
    if (ts.d) (void)hipFree(ts.d);
    ts.d = nullptr;
    ts.cap = 0;
    MFB_HIP(hipMalloc(&ts.d, sizeof(int2) * (size_t)nt));
    ts.cap = (size_t)nt;
  }
  long long* doff = nullptr;
  int rc0 = mesh_ws(m, MW_OFF, sizeof(long long) * 2 * ne * S, &doff);
  if (rc0) return rc0;
  hipError_t he = hipMemcpyAsync(doff, off.data(), sizeof(long long) * 2 * ne * S,
                                 hipMemcpyHostToDevice, s);
  if (he == hipSuccess) {
    hipLaunchKernelGGL(k_fill_lists, dim3(ne), dim3(256), 0, s, d, cls, doff, lists);
    if (nt) hipLaunchKernelGGL(k_fill_tiles, dim3(ne), dim3(64), 0, s, d, counts, doff,
                               doff + (size_t)ne * S, ts.d);
    he = hipGetLastError();
  }
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  if (he != hipSuccess) {
    set_error("build_lists: %s", hipGetErrorString(he));
    return MFB_ERR_HIP;
  }
  regular_points = 0;
  for (int r = 0; r < d.n_ps; r++) regular_points += cls_cnt[r] * d.ps_gln[r] * d.ps_gln[r];
  return MFB_OK;
}

// collocation points of a mesh (host copies; see DevMesh)
struct CollocSet {
  int ncp = 0;
  std::vector<double> xc, cxi;
  std::vector<int> crow, cel;
  std::vector<unsigned char> cshared, nmca;
};

static const double kXiQ9H[9][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}, {0, -1},
                                    {1, 0},   {0, 1},  {-1, 0}, {0, 0}};
static const int kEdgeQ9H[4][3] = {{0, 1, 4}, {1, 2, 5}, {2, 3, 6}, {3, 0, 7}};

// The reference's default BEM formulation of an ordinary boundary
// (src/assign_default_bem_formulation.f90:74-93): nodal SBIE on every node, except the nodes on
// the boundary of the boundary -- element edges that no other element shares
// (lib/fbem/src/data_structures.f90:1620-1640) -- which take SBIE-MCA: one collocation point in
// every element holding the node, at xi_node (1 - delta) (fbem_move_xi1xi2_from_edge,
// lib/fbem/src/geometry.f90:5631-5658; src/build_data_at_collocation_points.f90:125-147), x_i the
// element's position there.  Points are listed node by node; an MCA node's points follow its
// incidences in element order.
static void default_collocation(int nn, int ne, const double* x, const int* conn,
                                const std::vector<int>& ip, const std::vector<int>& ie,
                                const std::vector<int>& ik, double delta, CollocSet& cs) {
  std::map<std::tuple<int, int, int>, int> edges;
  for (int e = 0; e < ne; e++)
    for (int k = 0; k < 4; k++) {
      const int a = conn[9 * e + kEdgeQ9H[k][0]], b = conn[9 * e + kEdgeQ9H[k][1]];
      edges[std::make_tuple(std::min(a, b), std::max(a, b), conn[9 * e + kEdgeQ9H[k][2]])]++;
    }
  cs.nmca.assign(nn, 0);
  for (int e = 0; e < ne; e++)
    for (int k = 0; k < 4; k++) {
      const int a = conn[9 * e + kEdgeQ9H[k][0]], b = conn[9 * e + kEdgeQ9H[k][1]];
      const int c = conn[9 * e + kEdgeQ9H[k][2]];
      if (edges[std::make_tuple(std::min(a, b), std::max(a, b), c)] == 1)
        cs.nmca[a] = cs.nmca[b] = cs.nmca[c] = 1;
    }
  cs.ncp = 0;
  for (int i = 0; i < nn; i++) {
    if (!cs.nmca[i]) {
      for (int c = 0; c < 3; c++) cs.xc.push_back(x[3 * i + c]);
      cs.cxi.push_back(0.0);
      cs.cxi.push_back(0.0);
      cs.crow.push_back(i);
      cs.cel.push_back(-1);
      cs.cshared.push_back(0);
      cs.ncp++;
      continue;
    }
    for (int q = ip[i]; q < ip[i + 1]; q++) {
      const int e = ie[q], k = ik[q];
      const double xi1 = kXiQ9H[k][0] * (1.0 - delta), xi2 = kXiQ9H[k][1] * (1.0 - delta);
      double X[27], xp[3];
      for (int n = 0; n < 9; n++)
        for (int c = 0; c < 3; c++) X[3 * n + c] = x[3 * conn[9 * e + n] + c];
      q9_x(X, xi1, xi2, xp);
      for (int c = 0; c < 3; c++) cs.xc.push_back(xp[c]);
      cs.cxi.push_back(xi1);
      cs.cxi.push_back(xi2);
      cs.crow.push_back(i);
      cs.cel.push_back(e);
      cs.cshared.push_back(1);
      cs.ncp++;
    }
  }
}

// explicit: a caller-given collocation set (batches of independent pairs); otherwise the
// reference's default formulation with MCA parameter mca_delta.  classify_all = false leaves
// classification and lists to the caller.
static int mesh_create_impl(int nn, int ne, const double* x, const int* conn, const int* ctype,
                            const double* cval, double qsi_relative_error, int qsi_ns_max,
                            int n_precalsets, const int* precalset_gln, double geometric_tolerance,
                            const CollocSet* explicit_cs, double mca_delta, bool classify_all,
                            mfb_mesh_t* out) {
  MFB_ARG(out && x && conn && ctype && cval && precalset_gln, "null argument");
  MFB_ARG(nn > 0 && ne > 0, "empty mesh");
  MFB_ARG(n_precalsets >= 1 && n_precalsets <= kMaxPrecalsets, "1 <= n_precalsets <= 16");
  MFB_ARG(qsi_relative_error >= 1e-15 && qsi_relative_error <= 1e-3,
          "qsi_relative_error must be between 1e-15 and 1e-3 (read_settings.f90:81)");
  MFB_ARG(qsi_ns_max >= 1 && qsi_ns_max <= kMaxNs,
          "qsi_ns_max must be between 1 and 20 (the reference's default is 16)");
  for (int r = 0; r < n_precalsets; r++) {
    // one precalculated rule's kernel values of a pair stay in the LDS of one workgroup
    MFB_ARG(precalset_gln[r] >= 2 && precalset_gln[r] <= kMaxPrecalGln,
            "precalculated rules must be of order 2..21 (the reference's default is 2..9)");
    if (r) MFB_ARG(precalset_gln[r] > precalset_gln[r - 1], "precalsets must be sorted and unique");
  }
  for (long long q = 0; q < 9LL * ne; q++) MFB_ARG(conn[q] >= 0 && conn[q] < nn, "connectivity out of range");
  for (long long q = 0; q < 3LL * nn; q++)
    MFB_ARG(ctype[q] == 0 || ctype[q] == 1 || ctype[q] == 10, "ctype must be 0, 1 or 10");
  int rc = upload_gl_tables();
  if (rc == MFB_OK) rc = upload_polar_table();
  if (rc) return rc;

  // node -> (element, local node) incidences, in element order
  std::vector<int> cnt(nn + 1, 0), ip(nn + 1, 0), ie(9LL * ne), ik(9LL * ne);
  for (long long q = 0; q < 9LL * ne; q++) cnt[conn[q]]++;
  for (int i = 0; i < nn; i++) ip[i + 1] = ip[i] + cnt[i];
  std::vector<int> cur(ip.begin(), ip.end() - 1);
  for (int e = 0; e < ne; e++)
    for (int k = 0; k < 9; k++) {
      int j = conn[9 * e + k];
      ie[cur[j]] = e;
      ik[cur[j]] = k;
      cur[j]++;
    }
  CollocSet dcs;
  if (!explicit_cs) default_collocation(nn, ne, x, conn, ip, ie, ik, mca_delta, dcs);
  const CollocSet& cs = explicit_cs ? *explicit_cs : dcs;
  const int ncp = cs.ncp;
  MFB_ARG(ncp > 0 && (long long)ncp * ne < (1LL << 31),
          "collocation points x elements must fit in 31 bits");

  mfb_mesh_s* m = new mfb_mesh_s();
  DevMesh& d = m->dm;
  d.nn = nn;
  d.ne = ne;
  d.ncp = ncp;
  d.n_ps = n_precalsets;
  d.ns_max = qsi_ns_max;
  d.qsi_re = qsi_relative_error;
  d.tol = geometric_tolerance;
  d.ps_off[0] = 0;
  for (int r = 0; r < n_precalsets; r++) {
    d.ps_gln[r] = precalset_gln[r];
    d.ps_off[r + 1] = d.ps_off[r] + precalset_gln[r] * precalset_gln[r];
  }
  d.npts = d.ps_off[n_precalsets];
  host_qsi_params(qsi_relative_error, &m->qsp);
  host_qsi_params(1e-15, &m->qsp_li);
  m->n_cls = n_precalsets + 2;
  m->h_crow = cs.crow;
  m->h_cel = cs.cel;
  m->h_xc = cs.xc;
  m->h_x.assign(x, x + 3LL * nn);
  m->h_cval.assign(cval, cval + 3LL * nn);
  m->h_conn.assign(conn, conn + 9LL * ne);
  m->h_ctype.assign(ctype, ctype + 3LL * nn);
  m->h_ps.assign(precalset_gln, precalset_gln + n_precalsets);
  m->h_tol = geometric_tolerance;
  m->n_mca = 0;
  for (unsigned char f : cs.nmca) m->n_mca += f;

  // greedy colouring over shared nodes, elements in index order
  std::vector<int> ecol(ne, -1);
  bool color_overflow = false;
  {
    constexpr int kMaxColors = 64;
    std::vector<int> seen(kMaxColors, -1);
    for (int e = 0; e < ne; e++) {
      for (int k = 0; k < 9; k++) {
        const int j = conn[9 * e + k];
        for (int q = ip[j]; q < ip[j + 1]; q++)
          if (ecol[ie[q]] >= 0) seen[ecol[ie[q]]] = e;
      }
      int c = 0;
      while (c < kMaxColors && seen[c] == e) c++;
      if (c == kMaxColors) { color_overflow = true; c = 0; }
      ecol[e] = c;
      m->n_colors = std::max(m->n_colors, c + 1);
    }
  }
  std::vector<int> eorder(ne);
  for (int e = 0; e < ne; e++) eorder[e] = e;
  std::stable_sort(eorder.begin(), eorder.end(), [&](int a, int b) { return ecol[a] < ecol[b]; });

  auto fail = [&](int code) {
    mesh_release(m);
    delete m;
    return code;
  };
  if (color_overflow) {
    set_error("mfb_mesh_create: more than 64 elements share nodes around one element");
    return fail(MFB_ERR_GEOMETRY);
  }
#define TRY(call)               \
  do {                          \
    int rc_ = (call);           \
    if (rc_) return fail(rc_);  \
  } while (0)
#define TRYH(call)                                                                 \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      set_error("%s failed: %s", #call, hipGetErrorString(e_));                    \
      return fail(MFB_ERR_HIP);                                                    \
    }                                                                              \
  } while (0)
  TRY(dalloc(&m->x, 3LL * nn));
  TRY(dalloc(&m->cval, 3LL * nn));
  TRY(dalloc(&m->ctype, 3LL * nn));
  TRY(dalloc(&m->conn, 9LL * ne));
  TRY(dalloc(&m->elem, 8LL * ne));
  TRY(dalloc(&m->ps, (size_t)ne * d.npts * kPsStride));
  TRY(dalloc(&m->nfn, 3LL * nn));
  TRY(dalloc(&m->inc_ptr, nn + 1));
  TRY(dalloc(&m->inc_e, 9LL * ne));
  TRY(dalloc(&m->inc_k, 9LL * ne));
  TRY(dalloc(&m->xc, 3LL * ncp));
  TRY(dalloc(&m->cxi, 2LL * ncp));
  TRY(dalloc(&m->crow, ncp));
  TRY(dalloc(&m->cel, ncp));
  TRY(dalloc(&m->cshared, ncp));
  TRY(dalloc(&m->nmca, nn));
  TRY(dalloc(&m->err, 1));
  TRY(dalloc(&m->counters, 2));
  TRY(dalloc(&m->d_qsp, 1));
  TRY(dalloc(&m->d_qsp_li, 1));
  TRY(dalloc(&m->d_kp, 1));
  TRYH(hipMemcpy(m->d_qsp, &m->qsp, sizeof(QsiParams), hipMemcpyHostToDevice));
  for (hipEvent_t& e : m->ev) TRYH(hipEventCreate(&e));
  TRYH(hipMemcpy(m->d_qsp_li, &m->qsp_li, sizeof(QsiParams), hipMemcpyHostToDevice));
  TRYH(hipMemcpy(m->x, x, sizeof(double) * 3 * nn, hipMemcpyHostToDevice));
  TRYH(hipMemcpy(m->cval, cval, sizeof(double) * 3 * nn, hipMemcpyHostToDevice));
  TRYH(hipMemcpy(m->ctype, ctype, sizeof(int) * 3 * nn, hipMemcpyHostToDevice));
  TRYH(hipMemcpy(m->conn, conn, sizeof(int) * 9 * ne, hipMemcpyHostToDevice));
  TRYH(hipMemcpy(m->inc_ptr, ip.data(), sizeof(int) * (nn + 1), hipMemcpyHostToDevice));
  TRYH(hipMemcpy(m->inc_e, ie.data(), sizeof(int) * 9 * ne, hipMemcpyHostToDevice));
  TRYH(hipMemcpy(m->inc_k, ik.data(), sizeof(int) * 9 * ne, hipMemcpyHostToDevice));
  TRYH(hipMemcpy(m->xc, cs.xc.data(), sizeof(double) * 3 * ncp, hipMemcpyHostToDevice));
  TRYH(hipMemcpy(m->cxi, cs.cxi.data(), sizeof(double) * 2 * ncp, hipMemcpyHostToDevice));
  TRYH(hipMemcpy(m->crow, cs.crow.data(), sizeof(int) * ncp, hipMemcpyHostToDevice));
  TRYH(hipMemcpy(m->cel, cs.cel.data(), sizeof(int) * ncp, hipMemcpyHostToDevice));
  TRYH(hipMemcpy(m->cshared, cs.cshared.data(), ncp, hipMemcpyHostToDevice));
  if (cs.nmca.empty()) TRYH(hipMemset(m->nmca, 0, nn));
  else TRYH(hipMemcpy(m->nmca, cs.nmca.data(), nn, hipMemcpyHostToDevice));
  TRYH(hipMemset(m->err, 0, sizeof(int)));
  TRYH(hipMemset(m->counters, 0, 2 * sizeof(long long)));
  d.x = m->x; d.conn = m->conn; d.ctype = m->ctype; d.cval = m->cval; d.elem = m->elem;
  d.ps = m->ps; d.nfn = m->nfn; d.inc_ptr = m->inc_ptr; d.inc_e = m->inc_e; d.inc_k = m->inc_k;
  d.xc = m->xc; d.cxi = m->cxi; d.crow = m->crow; d.cel = m->cel; d.cshared = m->cshared;
  d.nmca = m->nmca;

  hipLaunchKernelGGL(k_elem_setup, dim3((ne + 63) / 64), dim3(64), 0, 0, d, m->elem);
  TRYH(hipGetLastError());
  long long tot = (long long)ne * d.npts;
  hipLaunchKernelGGL(k_precalc, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0, d, m->ps);
  TRYH(hipGetLastError());
  hipLaunchKernelGGL(k_node_normals, dim3((nn + 127) / 128), dim3(128), 0, 0, d, m->nfn);
  TRYH(hipGetLastError());
  m->ecol = ecol;
  m->eorder = eorder;
  m->tiles.off.assign(m->n_colors, 0);
  m->tiles.cnt.assign(m->n_colors, 0);
  m->cls_off.assign(m->n_cls + 1, 0);
  m->cls_cnt.assign(m->n_cls, 0);
  if (classify_all) {
    int* counts = nullptr;
    const int S = kMaxPrecalsets + 2;
    TRY(dalloc(&m->cls, (size_t)ncp * ne));
    TRY(dalloc(&counts, (size_t)ne * S));
    TRYH(hipMemset(counts, 0, sizeof(int) * ne * S));
    hipLaunchKernelGGL(k_classify, dim3(ne), dim3(256), 0, 0, d, m->d_qsp, m->cls, counts);
    TRYH(hipGetLastError());
    TRY(dalloc(&m->lists, (size_t)ncp * ne));
    {
      const int rc = build_lists(m, m->cls, counts, m->lists, m->cls_off, m->cls_cnt, m->tiles,
                                 m->regular_points, 0);
      (void)hipFree(counts);
      if (rc) return fail(rc);
    }
    // decision checksum (host) and regular-rule quadrature points
    std::vector<unsigned char> hcls((size_t)ncp * ne);
    TRYH(hipMemcpy(hcls.data(), m->cls, hcls.size(), hipMemcpyDeviceToHost));
    long long s = 0;
    for (size_t q = 0; q < hcls.size(); q++) s += (long long)(hcls[q] + 1) * (long long)(q % 97 + 1);
    m->decisions = s;
  }
#undef TRY
#undef TRYH
  *out = m;
  return MFB_OK;
}

extern "C" int mfb_mesh_create(int nn, int ne, const double* x, const int* conn, const int* ctype,
                               const double* cval, double qsi_relative_error, int qsi_ns_max,
                               int n_precalsets, const int* precalset_gln,
                               double geometric_tolerance, mfb_mesh_t* out) {
  // default_mca_boundary_delta (src/assign_default_bem_formulation.f90:40)
  return mesh_create_impl(nn, ne, x, conn, ctype, cval, qsi_relative_error, qsi_ns_max,
                          n_precalsets, precalset_gln, geometric_tolerance, nullptr, 0.05, true, out);
}

extern "C" int mfb_mesh_node_normals(mfb_mesh_t m, double* nfn) {
  MFB_ARG(m && nfn, "null argument");
  MFB_HIP(hipMemcpy(nfn, m->nfn, sizeof(double) * 3 * m->dm.nn, hipMemcpyDeviceToHost));
  return MFB_OK;
}

extern "C" int mfb_mesh_collocation(mfb_mesh_t m, int* ncp, double* xc, int* crow, int* cel) {
  MFB_ARG(m && ncp, "null argument");
  *ncp = m->dm.ncp;
  if (xc) std::memcpy(xc, m->h_xc.data(), sizeof(double) * 3 * m->dm.ncp);
  if (crow) std::memcpy(crow, m->h_crow.data(), sizeof(int) * m->dm.ncp);
  if (cel) std::memcpy(cel, m->h_cel.data(), sizeof(int) * m->dm.ncp);
  return MFB_OK;
}

extern "C" int mfb_mesh_destroy(mfb_mesh_t m) {
  if (!m) return MFB_OK;
  mesh_release(m);
  delete m;
  return MFB_OK;
}

extern "C" int mfb_mesh_pair_classes(mfb_mesh_t m, unsigned char* classes) {
  MFB_ARG(m && classes, "null argument");
  MFB_ARG(m->cls, "this mesh holds no class array");
  MFB_HIP(hipMemcpy(classes, m->cls, (size_t)m->dm.ncp * m->dm.ne, hipMemcpyDeviceToHost));
  return MFB_OK;
}

extern "C" int mfb_mesh_learned_classes(mfb_mesh_t m, unsigned char* classes) {
  MFB_ARG(m && classes, "null argument");
  MFB_ARG(m->lcls && m->learned_last, "the last assembly of this mesh did not use a network");
  MFB_HIP(hipMemcpy(classes, m->lcls, (size_t)m->dm.ncp * m->dm.ne, hipMemcpyDeviceToHost));
  return MFB_OK;
}

extern "C" int mfb_mesh_stats(mfb_mesh_t m, mfb_assembly_stats* s) {
  MFB_ARG(m && s, "null argument");
  long long c[2];
  MFB_HIP(hipMemcpy(c, m->counters, sizeof c, hipMemcpyDeviceToHost));
  const int np = m->dm.n_ps;
  const bool L = m->learned_last;
  const std::vector<long long>& cnt = L ? m->lcls_cnt : m->cls_cnt;
  const long long rp = L ? m->l_regular_points : m->regular_points;
  s->pairs = (long long)m->dm.ncp * m->dm.ne;
  s->singular = cnt[np + 1];
  s->adaptive = cnt[np];
  s->pre = 0;
  for (int r = 0; r < np; r++) s->pre += cnt[r];
  s->points = rp + c[0];
  s->leaves = c[1];
  s->decisions = m->decisions;
  s->regular_points = rp;
  s->learned = L ? 1 : 0;
  s->judge_pairs = m->judge_n;
  s->judge_max_err = m->judge_max;
  s->judge_mean_err = m->judge_mean;
  s->judge_ref_pairs = m->judge_ref_n;
  s->judge_ref_max_err = m->judge_ref_max;
  s->judge_promoted = m->judge_promoted;
  s->judge_first_max_err = m->judge_promoted ? m->judge_first_max : m->judge_max;
  return MFB_OK;
}

template <bool H>
static int run_judge(mfb_mesh_t m, const KernelParams& kp, hipStream_t s);

// k_regular of a rule above 64 KB of LDS (orders >= 14 harmonic, >= 18 static): raise the
// workgroup's dynamic LDS limit (up to the 160 KB of a gfx950 CU)
template <bool H>
static hipError_t set_kreg_lds(size_t lds) {
  hipError_t e = hipFuncSetAttribute((const void*)k_regular<H, kBlocks>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_regular<H, kVector>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  return e;
}

// Network classes of every pair for this call (k h from the wavenumber kw), then the lists.
static int learned_lists(mfb_mesh_t m, double kw, hipStream_t s) {
  const DevMesh& d = m->dm;
  const long long np = (long long)d.ne * d.ncp;
  hipLaunchKernelGGL(k_set_kh, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, d, m->feats, kw);
  int rc = mfb_qnet_classes_device(m->qnet, m->feats, (int)np, m->qconf, m->qcls, s);
  if (rc) return rc;
  int* counts = nullptr;
  rc = mesh_ws(m, MW_COUNTS, sizeof(int) * (size_t)d.ne * (kMaxPrecalsets + 2), &counts);
  if (rc) return rc;
  hipLaunchKernelGGL(k_learned_map, dim3(d.ne), dim3(256), 0, s, d, m->cls, m->qcls, m->d_promo,
                     m->lcls, counts);
  hipError_t he = hipGetLastError();
  if (he != hipSuccess) {
    set_error("k_learned_map: %s", hipGetErrorString(he));
    return MFB_ERR_HIP;
  }
  return build_lists(m, m->lcls, counts, m->llists, m->lcls_off, m->lcls_cnt, m->ltiles,
                     m->l_regular_points, s);
}

// The kernels of one whole-boundary assembly into A, b (scatter mode) on stream s, reading the
// kernel constants at kpp (device): zeroing, the precalculated-rule tiles (one launch per element
// colour), adaptive and singular pairs, free terms.  No host synchronisation and no memset: the
// sequence can be captured into a graph.  events: record the mesh's timing events around the
// kernel groups (not while capturing).
template <bool H>
static int enqueue_scatter(mfb_mesh_t m, const KernelParams* kpp, bool learned, double nu,
                           typename Val<H>::T* A, long long lda, typename Val<H>::T* b,
                           hipStream_t s, bool events) {
  typedef typename Val<H>::T T;
  const DevMesh& d = m->dm;
  const int n = 3 * d.nn;
  int* lists = learned ? m->llists : m->lists;
  const std::vector<long long>& cls_off = learned ? m->lcls_off : m->cls_off;
  const std::vector<long long>& cls_cnt = learned ? m->lcls_cnt : m->cls_cnt;
  const TileSet& ts = learned ? m->ltiles : m->tiles;
  if (events) MFB_HIP(hipEventRecord(m->ev[0], s));
  {
    // A, b := 0: the runtime's fill (6.1 TB/s at 30k) on a stream that is not being captured;
    // under capture (the batch graphs) a kernel, since memset nodes are kept out of graphs
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    MFB_HIP(hipStreamIsCapturing(s, &cap));
    if (cap == hipStreamCaptureStatusNone && lda == n) {
      MFB_HIP(hipMemsetAsync(A, 0, sizeof(T) * (size_t)n * n, s));
      MFB_HIP(hipMemsetAsync(b, 0, sizeof(T) * (size_t)n, s));
    } else {
      long long tot = (long long)n * n;
      long long nblk = std::min<long long>((tot + 255) / 256, 1 << 14);   // 5.6 TB/s; 1 << 16: 5.1
      nblk = std::max<long long>(nblk, (n + 255) / 256);
      hipLaunchKernelGGL(k_zero<T>, dim3((unsigned)nblk), dim3(256), 0, s, A, lda, n, b);
    }
  }
  if (events) MFB_HIP(hipEventRecord(m->ev[1], s));
  // precalculated rules: one launch per element colour over its tiles (elements of one colour
  // share no node: the pairs of a launch write disjoint entries of A, a plain read-modify-write;
  // the right-hand side keeps atomics)
  if (events) m->last_regular_launches = 0;
  for (int k = 0; k < m->n_colors; k++) {
    if (!ts.cnt[k]) continue;
    if (events) m->last_regular_launches++;
    hipLaunchKernelGGL((k_regular_tile<H>), dim3((unsigned)ts.cnt[k]), dim3(kTileThreads), 0, s, d,
                       kpp, lists, ts.d + ts.off[k], A, lda, b);
  }
  if (events) MFB_HIP(hipEventRecord(m->ev[2], s));
  if (cls_cnt[d.n_ps])
    hipLaunchKernelGGL(k_adaptive<H>, dim3((unsigned)cls_cnt[d.n_ps]), dim3(kWave), 0, s, d,
                       m->d_qsp, kpp, lists + cls_off[d.n_ps], (int)cls_cnt[d.n_ps], A, lda, b,
                       m->counters, m->err, (T*)nullptr, cls_off[d.n_ps], VecOut<T>{});
  if (events) MFB_HIP(hipEventRecord(m->ev[3], s));
  if (cls_cnt[d.n_ps + 1])
    hipLaunchKernelGGL(k_singular<H>, dim3((unsigned)cls_cnt[d.n_ps + 1]), dim3(kWave), 0, s, d,
                       m->d_qsp_li, kpp, lists + cls_off[d.n_ps + 1], (int)cls_cnt[d.n_ps + 1], A,
                       lda, b, m->counters, (T*)nullptr, cls_off[d.n_ps + 1], VecOut<T>{});
  if (events) MFB_HIP(hipEventRecord(m->ev[4], s));
  hipLaunchKernelGGL(k_freeterm<T>, dim3((d.nn + 127) / 128), dim3(128), 0, s, d, nu, A, lda, b,
                     m->err);
  if (m->n_mca)
    hipLaunchKernelGGL(k_freeterm_mca<T>, dim3((d.ncp + 127) / 128), dim3(128), 0, s, d, A, lda, b);
  if (events) MFB_HIP(hipEventRecord(m->ev[5], s));
  MFB_HIP(hipGetLastError());
  return MFB_OK;
}

// blocks == nullptr, vec.out == nullptr: assemble A, b.  blocks: write the 9x3x3 h and g blocks
// of every listed pair to blocks[162*q ...], q = position in the class lists (see
// mfb_pair_blocks).  vec.out: the interior displacements of the collocation points (vec.out is
// overwritten).  Both skip the zeroing of A, the scatter and the free terms.  kw: shear
// wavenumber (0 static), a network feature.
template <bool H>
static int assemble(mfb_mesh_t m, const KernelParams& kp, double nu, double* Araw, long long lda,
                    double* braw, void* stream, double* blocks_raw = nullptr, double kw = 0.0,
                    VecOut<typename Val<H>::T> vec = VecOut<typename Val<H>::T>{}) {
  typedef typename Val<H>::T T;
  T* blocks = (T*)blocks_raw;
  const bool vmode = vec.out != nullptr;
  MFB_ARG(m && ((Araw && braw) || blocks || vmode), "null argument");
  const DevMesh& d = m->dm;
  const int n = 3 * d.nn;
  MFB_ARG(blocks || vmode || lda >= n, "lda < n_dof");
  hipStream_t s = (hipStream_t)stream;
  T* A = (T*)Araw;
  T* b = (T*)braw;
  // kernel constants of this call (ordered before the launches on the same stream)
  MFB_HIP(hipMemcpyAsync(m->d_kp, &kp, sizeof(KernelParams), hipMemcpyHostToDevice, s));
  MFB_HIP(hipStreamSynchronize(s));
  // rule choice: the reference's estimator (classes fixed at mesh creation) or the network
  const bool learned = m->qnet && !blocks && !vmode;
  m->learned_last = learned;
  if (learned) {
    int rc = learned_lists(m, kw, s);
    if (rc) return rc;
  }
  int* lists = learned ? m->llists : m->lists;
  const std::vector<long long>& cls_off = learned ? m->lcls_off : m->cls_off;
  const std::vector<long long>& cls_cnt = learned ? m->lcls_cnt : m->cls_cnt;
  // counters and error flag are cleared by a kernel (no memset nodes if the caller captures)
  if (!blocks && !vmode) {
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, m->counters, m->err);
    const int rc = enqueue_scatter<H>(m, m->d_kp, learned, nu, A, lda, b, s, true);
    if (rc) return rc;
  } else {
    MFB_HIP(hipEventRecord(m->ev[0], s));
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, m->counters, m->err);
    if (vmode)
      hipLaunchKernelGGL(k_fill_zero<T>, dim3((3 * d.ncp + 255) / 256), dim3(256), 0, s, vec.out,
                         3LL * d.ncp);
    MFB_HIP(hipEventRecord(m->ev[1], s));
    m->last_regular_launches = 0;
    // per-pair blocks and interior displacements have no shared entries: one launch per rule
    for (int r = 0; r < d.n_ps; r++) {
      const long long cnt = cls_cnt[r], o = cls_off[r];
      if (!cnt) continue;
      m->last_regular_launches++;
      int P = d.ps_gln[r] * d.ps_gln[r], B = P >= kWave ? 1 : kWave / P;
      long long grid = (cnt + B - 1) / B;
      const size_t lds = kreg_lds_bytes<H>(B * P);
      if (lds > 64 * 1024) MFB_HIP(set_kreg_lds<H>(lds));
      if (vmode)
        hipLaunchKernelGGL((k_regular<H, kVector>), dim3((unsigned)grid), dim3(kWave), lds, s, d,
                           m->d_kp, lists + o, (int)cnt, r, blocks, o, vec);
      else
        hipLaunchKernelGGL((k_regular<H, kBlocks>), dim3((unsigned)grid), dim3(kWave), lds, s, d,
                           m->d_kp, lists + o, (int)cnt, r, blocks, o, vec);
    }
    MFB_HIP(hipEventRecord(m->ev[2], s));
    if (cls_cnt[d.n_ps]) {
      hipLaunchKernelGGL(k_adaptive<H>, dim3((unsigned)cls_cnt[d.n_ps]), dim3(kWave), 0, s, d,
                         m->d_qsp, m->d_kp, lists + cls_off[d.n_ps], (int)cls_cnt[d.n_ps], A, lda,
                         b, m->counters, m->err, blocks, cls_off[d.n_ps], vec);
    }
    MFB_HIP(hipEventRecord(m->ev[3], s));
    if (cls_cnt[d.n_ps + 1]) {
      hipLaunchKernelGGL(k_singular<H>, dim3((unsigned)cls_cnt[d.n_ps + 1]), dim3(kWave), 0, s, d,
                         m->d_qsp_li, m->d_kp, lists + cls_off[d.n_ps + 1],
                         (int)cls_cnt[d.n_ps + 1], A, lda, b, m->counters, blocks,
                         cls_off[d.n_ps + 1], vec);
    }
    MFB_HIP(hipEventRecord(m->ev[4], s));
    MFB_HIP(hipEventRecord(m->ev[5], s));
  }
  MFB_HIP(hipGetLastError());
  int herr = 0;
  MFB_HIP(hipMemcpyAsync(&herr, m->err, sizeof(int), hipMemcpyDeviceToHost, s));
  MFB_HIP(hipStreamSynchronize(s));
  {
    float t[5];
    for (int q = 0; q < 5; q++) MFB_HIP(hipEventElapsedTime(&t[q], m->ev[q], m->ev[q + 1]));
    m->last_ms[0] = t[1];
    m->last_ms[1] = t[2];
    m->last_ms[2] = t[3];
    m->last_ms[3] = t[0] + t[4];
  }
  if (herr == MFB_ERR_LIMIT) {
    set_error("a free-term node has more incident elements than %d", kMaxIncid - 2);
    return MFB_ERR_LIMIT;
  }
  if (herr == MFB_ERR_GEOMETRY) {
    set_error("the normals/tangents configuration of a free-term node is not valid");
    return MFB_ERR_GEOMETRY;
  }
  if (learned && m->judge_samples > 0) {
    if (!m->in_redo) m->judge_promoted = 0;
    const int rc = run_judge<H>(m, kp, s);
    if (rc || !m->judge_promoted || m->in_redo) return rc;
    // rules were promoted: assemble once more with them (the judge then reports the new choice)
    const double first = m->judge_max;
    m->in_redo = true;
    const int rc2 = assemble<H>(m, kp, nu, Araw, lda, braw, stream, nullptr, kw);
    m->in_redo = false;
    m->judge_first_max = first;
    return rc2;
  }
  m->judge_n = 0;
  m->judge_ref_n = 0;
  return MFB_OK;
}

extern "C" int mfb_mesh_timings(mfb_mesh_t m, double* ms, int* regular_launches) {
  MFB_ARG(m && ms, "null argument");
  for (int q = 0; q < 4; q++) ms[q] = m->last_ms[q];
  if (regular_launches) *regular_launches = m->last_regular_launches;
  return MFB_OK;
}

static KernelParams static_params(double mu, double nu) {
  KernelParams kp;
  std::memset(&kp, 0, sizeof kp);
  kp.harmonic = 0;
  kp.cteu1 = 1.0 / (16.0 * kPi * mu * (1.0 - nu));
  kp.cteu2 = 3.0 - 4.0 * nu;
  kp.ctet1 = -1.0 / (8.0 * kPi * (1.0 - nu));
  kp.ctet2 = 1.0 - 2.0 * nu;
  return kp;
}

extern "C" int mfb_assemble_staela(mfb_mesh_t m, double mu, double nu, double* A, long long lda,
                                   double* b, void* stream) {
  MFB_ARG(mu > 0 && nu < 0.5 && nu > -1.0, "invalid elastic constants");
  return assemble<false>(m, static_params(mu, nu), nu, A, lda, b, stream);
}

extern "C" int mfb_assemble_harela(mfb_mesh_t m, double omega, double lambda_re, double lambda_im,
                                   double mu_re, double mu_im, double rho, double nu, double* A,
                                   long long lda, double* b, void* stream) {
  MFB_ARG(omega > 0 && rho > 0 && mu_re > 0, "invalid harmonic parameters (omega > 0 required)");
  KernelParams kp;
  std::memset(&kp, 0, sizeof kp);
  kp.harmonic = 1;
  host_har_params(lambda_re, lambda_im, mu_re, mu_im, rho, omega, &kp.h);
  return assemble<true>(m, kp, nu, A, lda, b, stream, nullptr, omega * std::sqrt(rho / mu_re));
}

// ------------------------------------------------------------------------------------------------
// Batches of assemblies: many (mesh, material, frequency) jobs enqueued back to back with no host
// synchronisation, and the same sequence captured once into a HIP graph and replayed (with new
// materials / frequencies written into the graph's parameter buffer between launches).  The rule
// choice is the reference's (classes and tiles fixed at mesh creation), so the launch sequence of a
// job does not depend on its frequency.  Every job reads its own kernel constants from one device
// array uploaded by a single host-to-device copy at the start of the sequence.
// ------------------------------------------------------------------------------------------------
static int job_params(const mfb_assembly_job& j, KernelParams* kp) {
  MFB_ARG(j.mesh && j.A && j.b && j.lda >= 3LL * j.mesh->dm.nn, "invalid assembly job");
  MFB_ARG(!j.mesh->qnet, "batched assembly uses the reference rule choice (mesh has a network)");
  if (j.harmonic) {
    MFB_ARG(j.omega > 0 && j.rho > 0 && j.mu_re > 0, "invalid harmonic job parameters");
    std::memset(kp, 0, sizeof *kp);
    kp->harmonic = 1;
    host_har_params(j.lambda_re, j.lambda_im, j.mu_re, j.mu_im, j.rho, j.omega, &kp->h);
  } else {
    MFB_ARG(j.mu_re > 0 && j.nu < 0.5 && j.nu > -1.0, "invalid static job parameters");
    *kp = static_params(j.mu_re, j.nu);
  }
  return MFB_OK;
}

// the jobs' kernels on s: one parameter upload (host_kp pinned), a counter / error reset per
// distinct mesh, then every job's sequence
static int enqueue_batch(int njobs, const mfb_assembly_job* jobs, const KernelParams* host_kp,
                         KernelParams* dev_kp, hipStream_t s) {
  MFB_HIP(hipMemcpyAsync(dev_kp, host_kp, sizeof(KernelParams) * njobs, hipMemcpyHostToDevice, s));
  std::vector<mfb_mesh_t> seen;
  for (int q = 0; q < njobs; q++) {
    mfb_mesh_t m = jobs[q].mesh;
    if (std::find(seen.begin(), seen.end(), m) != seen.end()) continue;
    seen.push_back(m);
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, m->counters, m->err);
  }
  for (int q = 0; q < njobs; q++) {
    const mfb_assembly_job& j = jobs[q];
    const int rc = j.harmonic
        ? enqueue_scatter<true>(j.mesh, dev_kp + q, false, j.nu, (zd*)j.A, j.lda, (zd*)j.b, s, false)
        : enqueue_scatter<false>(j.mesh, dev_kp + q, false, j.nu, j.A, j.lda, j.b, s, false);
    if (rc) return rc;
  }
  return MFB_OK;
}

// the error flags the kernels of the meshes' last sequences raised (after a synchronisation)
static int batch_errors(int njobs, const mfb_assembly_job* jobs) {
  for (int q = 0; q < njobs; q++) {
    int herr = 0;
    MFB_HIP(hipMemcpy(&herr, jobs[q].mesh->err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr == MFB_ERR_LIMIT) {
      set_error("job %d: a free-term node has more incident elements than %d", q, kMaxIncid - 2);
      return MFB_ERR_LIMIT;
    }
    if (herr == MFB_ERR_GEOMETRY) {
      set_error("job %d: the normals/tangents configuration of a free-term node is not valid", q);
      return MFB_ERR_GEOMETRY;
    }
  }
  return MFB_OK;
}

struct mfb_batch_graph_s {
  std::vector<mfb_assembly_job> jobs;
  KernelParams* host_kp = nullptr;   // pinned: the graph's copy node reads it at every launch
  KernelParams* dev_kp = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap = nullptr;
};

static void batch_graph_release(mfb_batch_graph_s* g) {
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->cap) (void)hipStreamDestroy(g->cap);
  if (g->host_kp) (void)hipHostFree(g->host_kp);
  if (g->dev_kp) (void)hipFree(g->dev_kp);
  delete g;
}

extern "C" int mfb_assemble_batch(int njobs, const mfb_assembly_job* jobs, void* stream) {
  MFB_ARG(njobs > 0 && jobs, "null argument");
  hipStream_t s = (hipStream_t)stream;
  KernelParams *hk = nullptr, *dk = nullptr;
  MFB_HIP(hipHostMalloc((void**)&hk, sizeof(KernelParams) * njobs));
  struct Free {
    KernelParams *h, *d;
    ~Free() { if (h) (void)hipHostFree(h); if (d) (void)hipFree(d); }
  } fr{hk, nullptr};
  for (int q = 0; q < njobs; q++) {
    const int rc = job_params(jobs[q], hk + q);
    if (rc) return rc;
  }
  MFB_HIP(hipMalloc(&dk, sizeof(KernelParams) * njobs));
  fr.d = dk;
  int rc = enqueue_batch(njobs, jobs, hk, dk, s);
  if (rc) return rc;
  MFB_HIP(hipStreamSynchronize(s));
  return batch_errors(njobs, jobs);
}

extern "C" int mfb_batch_graph_create(int njobs, const mfb_assembly_job* jobs,
                                      mfb_batch_graph_t* out) {
  MFB_ARG(njobs > 0 && jobs && out, "null argument");
  auto* g = new mfb_batch_graph_s();
  auto fail = [&](int rc) { batch_graph_release(g); return rc; };
  g->jobs.assign(jobs, jobs + njobs);
  if (hipHostMalloc((void**)&g->host_kp, sizeof(KernelParams) * njobs) != hipSuccess ||
      hipMalloc(&g->dev_kp, sizeof(KernelParams) * njobs) != hipSuccess ||
      hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking) != hipSuccess) {
    set_error("mfb_batch_graph_create: allocation failed");
    return fail(MFB_ERR_HIP);
  }
  for (int q = 0; q < njobs; q++) {
    const int rc = job_params(jobs[q], g->host_kp + q);
    if (rc) return fail(rc);
  }
  hipError_t he = hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal);
  if (he != hipSuccess) {
    set_error("hipStreamBeginCapture: %s", hipGetErrorString(he));
    return fail(MFB_ERR_HIP);
  }
  const int rc = enqueue_batch(njobs, jobs, g->host_kp, g->dev_kp, g->cap);
  he = hipStreamEndCapture(g->cap, &g->graph);
  if (rc) return fail(rc);
  if (he == hipSuccess) he = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (he != hipSuccess) {
    set_error("graph capture: %s", hipGetErrorString(he));
    return fail(MFB_ERR_HIP);
  }
  *out = g;
  return MFB_OK;
}

extern "C" int mfb_batch_graph_set_params(mfb_batch_graph_t g, int njobs,
                                          const mfb_assembly_job* jobs) {
  MFB_ARG(g && jobs && njobs == (int)g->jobs.size(), "job count differs from the graph's");
  for (int q = 0; q < njobs; q++) {
    const mfb_assembly_job &a = jobs[q], &c = g->jobs[q];
    MFB_ARG(a.mesh == c.mesh && a.A == c.A && a.b == c.b && a.lda == c.lda &&
                a.harmonic == c.harmonic,
            "a job's mesh, buffers or kind differ from the captured graph's");
  }
  // the previous launch may still read the parameters: wait for it before rewriting them
  MFB_HIP(hipStreamSynchronize(g->cap));
  for (int q = 0; q < njobs; q++) {
    KernelParams kp;
    const int rc = job_params(jobs[q], &kp);
    if (rc) return rc;
    g->host_kp[q] = kp;
    g->jobs[q] = jobs[q];
  }
  return MFB_OK;
}

extern "C" int mfb_batch_graph_launch(mfb_batch_graph_t g, void* stream) {
  MFB_ARG(g, "null argument");
  hipStream_t s = stream ? (hipStream_t)stream : g->cap;
  MFB_HIP(hipGraphLaunch(g->exec, s));
  if (s != g->cap) {   // set_params waits on the capture stream: order it after this launch
    hipEvent_t ev;
    MFB_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    MFB_HIP(hipEventRecord(ev, s));
    MFB_HIP(hipStreamWaitEvent(g->cap, ev, 0));
    MFB_HIP(hipEventDestroy(ev));
  }
  return MFB_OK;
}

extern "C" int mfb_batch_graph_check(mfb_batch_graph_t g) {
  MFB_ARG(g, "null argument");
  MFB_HIP(hipStreamSynchronize(g->cap));
  MFB_HIP(hipDeviceSynchronize());
  return batch_errors((int)g->jobs.size(), g->jobs.data());
}

extern "C" int mfb_batch_graph_destroy(mfb_batch_graph_t g) {
  if (!g) return MFB_OK;
  (void)hipStreamSynchronize(g->cap);
  batch_graph_release(g);
  return MFB_OK;
}

// ------------------------------------------------------------------------------------------------
// Learned rule choice: features and labels of listed pairs (pair id = e * nn + i)
// ------------------------------------------------------------------------------------------------
__global__ void k_gather_cls(const unsigned char* cls, const int* pairs, int np, int* out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < np) out[q] = cls[pairs[q]];
}

__global__ void k_scatter_rows(const double* src, int np, int nr, const int* idx, double* dst) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)np * nr) return;
  const int q = (int)(t / nr), r = (int)(t % nr);
  dst[(long long)idx[q] * nr + r] = src[t];
}

__global__ void k_fill_double(double* a, long long n, double v) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) a[q] = v;
}

// err (np x n_ps, device): error of every precalculated rule of the listed (non-singular) pairs
// against the adaptive rule run with the estimator tables d_qsp_ref
template <bool H>
static int rule_errors(mfb_mesh_t m, const KernelParams& kp, const int* dlist, int np,
                       const QsiParams* d_qsp_ref, double* err, long long* counters, int* errflag,
                       hipStream_t s, bool keep_overflow = false) {
  typedef typename Val<H>::T T;
  const DevMesh& d = m->dm;
  T *ref = nullptr, *blk = nullptr;
  int rc = mesh_ws(m, MW_REF, sizeof(T) * 162 * (size_t)np, &ref);
  if (!rc) rc = mesh_ws(m, MW_BLK, sizeof(T) * 162 * (size_t)np, &blk);
  if (rc) return rc;
  hipError_t he = hipMemcpyAsync(m->d_kp, &kp, sizeof(KernelParams), hipMemcpyHostToDevice, s);
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  if (he == hipSuccess) {
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, counters, errflag);
    hipLaunchKernelGGL(k_adaptive<H>, dim3((unsigned)np), dim3(kWave), 0, s, d, d_qsp_ref, m->d_kp,
                       dlist, np, (T*)nullptr, 0LL, (T*)nullptr, counters, errflag, ref, 0LL,
                       VecOut<T>{});
    for (int r = 0; r < d.n_ps; r++) {
      const int P = d.ps_gln[r] * d.ps_gln[r], B = P >= kWave ? 1 : kWave / P;
      hipLaunchKernelGGL((k_regular<H, kBlocks>), dim3((unsigned)((np + B - 1) / B)), dim3(kWave),
                         kreg_lds_bytes<H>(B * P), s, d, m->d_kp, dlist, np, r, blk, 0LL,
                         VecOut<T>{});
      hipLaunchKernelGGL(k_block_err<T>, dim3((unsigned)np), dim3(kWave), 0, s, ref, blk, d.n_ps,
                         r, err);
    }
    he = hipGetLastError();
  }
  int herr = 0;
  if (he == hipSuccess) he = hipMemcpyAsync(&herr, errflag, sizeof(int), hipMemcpyDeviceToHost, s);
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  if (he != hipSuccess) {
    set_error("rule_errors: %s", hipGetErrorString(he));
    rc = MFB_ERR_HIP;
  } else if (herr && !keep_overflow) {
    set_error("rule_errors: adaptive subdivision stack exceeded");
    rc = MFB_ERR_LIMIT;
  }
  return rc;
}

// The judge of the learned rule choice: an evenly spaced sample of the pairs the network gave a
// precalculated rule is integrated with the adaptive rule at the judge's tolerance; the sample's
// largest and mean block errors of the rules actually used are kept for mfb_mesh_stats.
template <bool H>
static int run_judge(mfb_mesh_t m, const KernelParams& kp, hipStream_t s) {
  const DevMesh& d = m->dm;
  const long long total = m->lcls_off[d.n_ps];
  m->judge_n = 0;
  m->judge_max = 0;
  m->judge_mean = 0;
  m->judge_ref_n = 0;
  m->judge_ref_max = 0;
  if (total == 0) return MFB_OK;
  const int K = (int)std::min<long long>(m->judge_samples, total);
  std::vector<long long> pos(K);
  std::vector<int> rk(K);
  for (int k = 0; k < K; k++) {
    pos[k] = std::min(total - 1, (long long)(((double)k + 0.5) * (double)total / (double)K));
    int r = 0;
    while (r + 1 < d.n_ps && pos[k] >= m->lcls_off[r + 1]) r++;
    rk[k] = r;
  }
  long long* dpos = nullptr;
  int* dsample = nullptr;
  double* err = nullptr;
  int rc = mesh_ws(m, MW_POS, sizeof(long long) * K, &dpos);
  if (!rc) rc = mesh_ws(m, MW_SAMPLE, sizeof(int) * K, &dsample);
  if (!rc) rc = mesh_ws(m, MW_ERR, sizeof(double) * (size_t)K * d.n_ps, &err);
  if (rc) return rc;
  hipError_t he = hipMemcpyAsync(dpos, pos.data(), sizeof(long long) * K, hipMemcpyHostToDevice, s);
  if (he == hipSuccess) {
    hipLaunchKernelGGL(k_gather_list, dim3((K + 255) / 256), dim3(256), 0, s, m->llists, dpos, K,
                       dsample);
    rc = rule_errors<H>(m, kp, dsample, K, m->d_qsp_judge, err, m->judge_counters, m->judge_err, s,
                        true);
  } else {
    set_error("run_judge: %s", hipGetErrorString(he));
    rc = MFB_ERR_HIP;
  }
  // the reference estimator's rule for the same pairs (when it is a precalculated one)
  std::vector<int> pc(K, -1);
  int* dpc = nullptr;
  if (rc == MFB_OK) {
    rc = mesh_ws(m, MW_PC, sizeof(int) * K, &dpc);
    if (rc) return rc;
    {
      hipLaunchKernelGGL(k_gather_cls, dim3((K + 255) / 256), dim3(256), 0, s, m->cls, dsample, K, dpc);
      he = hipMemcpyAsync(pc.data(), dpc, sizeof(int) * K, hipMemcpyDeviceToHost, s);
    }
    if (he == hipSuccess) he = hipStreamSynchronize(s);
    if (he != hipSuccess) {
      set_error("run_judge: %s", hipGetErrorString(he));
      rc = MFB_ERR_HIP;
    }
  }
  if (rc == MFB_OK) {
    std::vector<double> he_(static_cast<size_t>(K) * d.n_ps);
    he = hipMemcpy(he_.data(), err, sizeof(double) * he_.size(), hipMemcpyDeviceToHost);
    if (he != hipSuccess) {
      set_error("run_judge: %s", hipGetErrorString(he));
      rc = MFB_ERR_HIP;
    } else {
      double mx = 0, sum = 0, pmx = 0;
      long long pn = 0, nk = 0;
      for (int k = 0; k < K; k++) {
        const double e = he_[(size_t)k * d.n_ps + rk[k]];
        if (e != e) continue;  // no reference for this pair (subdivision overflow)
        nk++;
        mx = std::max(mx, e);
        sum += e;
        if (pc[k] >= 0 && pc[k] < d.n_ps) {
          pmx = std::max(pmx, he_[(size_t)k * d.n_ps + pc[k]]);
          pn++;
        }
      }
      m->judge_n = nk;
      m->judge_max = mx;
      m->judge_mean = nk ? sum / nk : 0.0;
      m->judge_ref_n = pn;
      m->judge_ref_max = pmx;
      if (m->judge_promote && !m->in_redo) {
        // a rule whose judged pairs miss the tolerance: the cheapest longer rule that meets it on
        // all of them, else the adaptive rule; the promotion applies to the network classes that
        // map to the rule (promo composes with earlier promotions)
        const double tol = d.qsi_re;
        bool changed = false;
        for (int r = 0; r < d.n_ps; r++) {
          double worst = 0;
          bool any = false;
          for (int k = 0; k < K; k++) {
            const double e = he_[(size_t)k * d.n_ps + r];
            if (rk[k] != r || e != e) continue;
            any = true;
            worst = std::max(worst, e);
          }
          if (!any || worst <= tol) continue;
          int to = d.n_ps;   // adaptive
          for (int r2 = 0; r2 < d.n_ps && to == d.n_ps; r2++) {
            if (d.ps_gln[r2] <= d.ps_gln[r]) continue;
            bool ok = true;
            for (int k = 0; k < K && ok; k++) {
              const double e = he_[(size_t)k * d.n_ps + r2];
              if (rk[k] == r && e == e && e > tol) ok = false;
            }
            if (ok) to = r2;
          }
          for (int q = 0; q < d.n_ps; q++)
            if (m->promo[q] == r) m->promo[q] = to;
          m->judge_promoted++;
          changed = true;
        }
        if (changed) {
          he = hipMemcpy(m->d_promo, m->promo.data(), sizeof(int) * d.n_ps, hipMemcpyHostToDevice);
          if (he != hipSuccess) {
            set_error("run_judge: %s", hipGetErrorString(he));
            rc = MFB_ERR_HIP;
          }
        }
      }
    }
  }
  return rc;
}

extern "C" int mfb_mesh_set_qnet(mfb_mesh_t m, mfb_qnet_t q, float confidence,
                                 double judge_qsi_relative_error, int judge_samples) {
  MFB_ARG(m, "null mesh");
  if (!q) {
    m->qnet = nullptr;
    return MFB_OK;
  }
  MFB_ARG(confidence > 0.f && confidence <= 1.f, "confidence must be in (0, 1]");
  MFB_ARG(judge_samples >= 0 && judge_qsi_relative_error > 0 && judge_qsi_relative_error < 1,
          "judge_samples >= 0 and 0 < judge_qsi_relative_error < 1 required");
  const DevMesh& d = m->dm;
  const long long np = (long long)d.ne * d.ncp;
  MFB_ARG(m->cls, "this mesh holds no class array");
  if (!m->feats) {
    MFB_HIP(hipMalloc(&m->feats, sizeof(float) * MFB_QNET_INPUTS * (size_t)np));
    MFB_HIP(hipMalloc(&m->qcls, sizeof(int) * (size_t)np));
    MFB_HIP(hipMalloc(&m->lcls, (size_t)np));
    MFB_HIP(hipMalloc(&m->llists, sizeof(int) * (size_t)np));
    MFB_HIP(hipMalloc(&m->d_qsp_judge, sizeof(QsiParams)));
    MFB_HIP(hipMalloc(&m->judge_counters, 2 * sizeof(long long)));
    MFB_HIP(hipMalloc(&m->judge_err, sizeof(int)));
    MFB_HIP(hipMalloc(&m->d_promo, sizeof(int) * kMaxPrecalsets));
    hipLaunchKernelGGL(k_all_features, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, 0, d,
                       m->feats);
    MFB_HIP(hipGetLastError());
    MFB_HIP(hipDeviceSynchronize());
  }
  QsiParams hq;
  host_qsi_params(judge_qsi_relative_error, &hq);
  MFB_HIP(hipMemcpy(m->d_qsp_judge, &hq, sizeof(QsiParams), hipMemcpyHostToDevice));
  m->qnet = q;
  m->qconf = confidence;
  m->judge_samples = judge_samples;
  // a new network starts without promotions
  m->promo.resize(d.n_ps);
  for (int r = 0; r < d.n_ps; r++) m->promo[r] = r;
  MFB_HIP(hipMemcpy(m->d_promo, m->promo.data(), sizeof(int) * d.n_ps, hipMemcpyHostToDevice));
  return MFB_OK;
}

extern "C" int mfb_mesh_set_judge_action(mfb_mesh_t m, int promote) {
  MFB_ARG(m && m->qnet, "the mesh has no network (mfb_mesh_set_qnet first)");
  MFB_ARG(m->judge_samples > 0 || !promote, "the judge samples no pairs (judge_samples = 0)");
  m->judge_promote = promote ? 1 : 0;
  return MFB_OK;
}

template <bool H>
static int pair_labels(mfb_mesh_t m, const KernelParams& kp, const int* pairs, int np, double tau,
                       int* labels, double* errors, hipStream_t s) {
  MFB_ARG(m && pairs && labels && np > 0 && tau > 0, "bad argument");
  const DevMesh& d = m->dm;
  const int nr = d.n_ps;
  std::vector<int> hp(np), hc(np);
  int* dcls = nullptr;
  MFB_HIP(hipMalloc(&dcls, sizeof(int) * np));
  hipLaunchKernelGGL(k_gather_cls, dim3((np + 255) / 256), dim3(256), 0, s, m->cls, pairs, np, dcls);
  MFB_HIP(hipMemcpyAsync(hc.data(), dcls, sizeof(int) * np, hipMemcpyDeviceToHost, s));
  MFB_HIP(hipMemcpyAsync(hp.data(), pairs, sizeof(int) * np, hipMemcpyDeviceToHost, s));
  MFB_HIP(hipStreamSynchronize(s));
  (void)hipFree(dcls);
  std::vector<int> flist, fidx;
  for (int q = 0; q < np; q++)
    if (hc[q] != d.n_ps + 1) {
      flist.push_back(hp[q]);
      fidx.push_back(q);
    }
  const int nf = (int)flist.size();
  hipLaunchKernelGGL(k_fill_int, dim3((np + 255) / 256), dim3(256), 0, s, labels, np, -1);
  if (errors)
    hipLaunchKernelGGL(k_fill_double, dim3((unsigned)(((long long)np * nr + 255) / 256)), dim3(256),
                       0, s, errors, (long long)np * nr, -1.0);
  if (nf == 0) {
    MFB_HIP(hipStreamSynchronize(s));
    return MFB_OK;
  }
  int *dl = nullptr, *di = nullptr;
  double* err = nullptr;
  MFB_HIP(hipMalloc(&dl, sizeof(int) * nf));
  MFB_HIP(hipMalloc(&di, sizeof(int) * nf));
  MFB_HIP(hipMalloc(&err, sizeof(double) * (size_t)nf * nr));
  MFB_HIP(hipMemcpyAsync(dl, flist.data(), sizeof(int) * nf, hipMemcpyHostToDevice, s));
  MFB_HIP(hipMemcpyAsync(di, fidx.data(), sizeof(int) * nf, hipMemcpyHostToDevice, s));
  int rc = rule_errors<H>(m, kp, dl, nf, m->d_qsp, err, m->counters, m->err, s, true);
  if (rc == MFB_OK) {
    hipLaunchKernelGGL(k_rule_labels, dim3((nf + 255) / 256), dim3(256), 0, s, err, nf, nr, d, tau,
                       di, labels);
    if (errors)
      hipLaunchKernelGGL(k_scatter_rows, dim3((unsigned)(((long long)nf * nr + 255) / 256)),
                         dim3(256), 0, s, err, nf, nr, di, errors);
    hipError_t he = hipGetLastError();
    if (he == hipSuccess) he = hipStreamSynchronize(s);
    if (he != hipSuccess) {
      set_error("pair_labels: %s", hipGetErrorString(he));
      rc = MFB_ERR_HIP;
    }
  }
  (void)hipFree(dl);
  (void)hipFree(di);
  (void)hipFree(err);
  return rc;
}

extern "C" int mfb_mesh_pair_features(mfb_mesh_t m, const int* pairs, int np, double wavenumber,
                                      float* feats, void* stream) {
  MFB_ARG(m && pairs && feats && np > 0 && wavenumber >= 0, "bad argument");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_pair_features, dim3((np + 127) / 128), dim3(128), 0, s, m->dm, pairs, np,
                     wavenumber, feats);
  MFB_HIP(hipGetLastError());
  MFB_HIP(hipStreamSynchronize(s));
  return MFB_OK;
}

extern "C" int mfb_mesh_pair_labels_staela(mfb_mesh_t m, double mu, double nu, const int* pairs,
                                           int np, double tau, int* labels, double* errors,
                                           void* stream) {
  MFB_ARG(mu > 0 && nu < 0.5 && nu > -1.0, "invalid elastic constants");
  return pair_labels<false>(m, static_params(mu, nu), pairs, np, tau, labels, errors,
                            (hipStream_t)stream);
}

extern "C" int mfb_mesh_pair_labels_harela(mfb_mesh_t m, double omega, double lambda_re,
                                           double lambda_im, double mu_re, double mu_im,
                                           double rho, const int* pairs, int np, double tau,
                                           int* labels, double* errors, void* stream) {
  MFB_ARG(omega > 0 && rho > 0 && mu_re > 0, "invalid harmonic parameters (omega > 0 required)");
  KernelParams kp;
  std::memset(&kp, 0, sizeof kp);
  kp.harmonic = 1;
  host_har_params(lambda_re, lambda_im, mu_re, mu_im, rho, omega, &kp.h);
  return pair_labels<true>(m, kp, pairs, np, tau, labels, errors, (hipStream_t)stream);
}

// Host scatter of one pair block (src/assemble_bem_staela_equation.f90:69-98).
template <typename V>
static void scatter_host(const int* ctype, const double* cval, const double* nfn, int reversion,
                         int ncol, const int* en, const V* h, const V* g, V* A, long long lda,
                         V* b) {
  for (int il = 0; il < 3; il++) {
    long long row = 3LL * ncol + il;
    for (int ik = 0; ik < 3; ik++)
      for (int kn = 0; kn < 9; kn++) {
        int sn = en[kn];
        long long col = 3LL * sn + ik;
        int o = kn + 9 * il + 27 * ik;
        int ct = ctype[3 * sn + ik];
        double cv = cval[3 * sn + ik];
        switch (ct) {
          case 0:
            A[row + lda * col] -= g[o];
            b[row] -= h[o] * cv;
            break;
          case 1:
            A[row + lda * col] += h[o];
            b[row] += g[o] * cv;
            break;
          default:
            A[row + lda * col] += h[o];
            if (!reversion) b[row] += g[o] * cv * nfn[3 * sn + ik];
            else b[row] -= g[o] * cv * nfn[3 * sn + ik];
        }
      }
  }
}

extern "C" int mfb_assemble_bem_staela_equation(int nn, const int* ctype, const double* cval,
                                                const double* node_normal, int reversion,
                                                int node_col, const int* enodes, const double* h,
                                                const double* g, double* A, long long lda,
                                                double* b) {
  MFB_ARG(ctype && cval && node_normal && enodes && h && g && A && b, "null argument");
  MFB_ARG(node_col >= 0 && node_col < nn && lda >= 3LL * nn, "index out of range");
  for (int k = 0; k < 9; k++) MFB_ARG(enodes[k] >= 0 && enodes[k] < nn, "element node out of range");
  scatter_host<double>(ctype, cval, node_normal, reversion, node_col, enodes, h, g, A, lda, b);
  return MFB_OK;
}

extern "C" int mfb_assemble_bem_harela_equation(int nn, const int* ctype, const double* cval,
                                                const double* node_normal, int reversion,
                                                int node_col, const int* enodes, const double* h,
                                                const double* g, double* A, long long lda,
                                                double* b) {
  MFB_ARG(ctype && cval && node_normal && enodes && h && g && A && b, "null argument");
  MFB_ARG(node_col >= 0 && node_col < nn && lda >= 3LL * nn, "index out of range");
  for (int k = 0; k < 9; k++) MFB_ARG(enodes[k] >= 0 && enodes[k] < nn, "element node out of range");
  typedef std::complex<double> z;
  scatter_host<z>(ctype, cval, node_normal, reversion, node_col, enodes, (const z*)h, (const z*)g,
                  (z*)A, lda, (z*)b);
  return MFB_OK;
}

// The BE-ordinary scatter with the reference's own DOF tables (node%row, node%col, ctype, cvalue,
// n_fn of src/assemble_bem_{staela,harela}_equation.f90:69-113): rows[il] = equation of direction
// il of the collocation node; cols[6 kn + k] = column of unknown k of element node kn (k < 3:
// u_k, k >= 3: t_{k-3}; -1 where the node has none).  ctype 2/3 (local-axes conditions: u and t
// both unknown, harmonic only) put h in the u column and -g in the t column.
template <typename V>
static int scatter_tables(int n_el_nodes, const int* rows, const int* cols, const int* ctype,
                          const V* cval, const double* n_fn, int reversion, const V* h, const V* g,
                          V* A, long long lda, V* b, bool allow_local) {
  const int ne = n_el_nodes;
  for (int il = 0; il < 3; il++) {
    const long long row = rows[il];
    for (int ik = 0; ik < 3; ik++)
      for (int kn = 0; kn < ne; kn++) {
        const int o = kn + ne * il + 3 * ne * ik;
        const int ct = ctype[3 * kn + ik];
        const V cv = cval[3 * kn + ik];
        const long long cu = cols[6 * kn + ik], ct_col = cols[6 * kn + 3 + ik];
        switch (ct) {
          case 0:
            if (ct_col < 0) return MFB_ERR_ARG;
            A[row + lda * ct_col] -= g[o];
            b[row] -= h[o] * cv;
            break;
          case 1:
            if (cu < 0) return MFB_ERR_ARG;
            A[row + lda * cu] += h[o];
            b[row] += g[o] * cv;
            break;
          case 10:
            if (cu < 0) return MFB_ERR_ARG;
            A[row + lda * cu] += h[o];
            if (!reversion) b[row] += g[o] * cv * n_fn[3 * kn + ik];
            else b[row] -= g[o] * cv * n_fn[3 * kn + ik];
            break;
          case 2:
          case 3:
            if (!allow_local || cu < 0 || ct_col < 0) return MFB_ERR_ARG;
            A[row + lda * cu] += h[o];
            A[row + lda * ct_col] -= g[o];
            break;
          default:
            return MFB_ERR_ARG;
        }
      }
  }
  return MFB_OK;
}

extern "C" int mfb_scatter_bem_staela_equation(int n_el_nodes, const int* rows, const int* cols,
                                               const int* ctype, const double* cvalue,
                                               const double* n_fn, int reversion, const double* h,
                                               const double* g, double* A, long long lda,
                                               double* b) {
  MFB_ARG(n_el_nodes > 0 && n_el_nodes <= 9 && rows && cols && ctype && cvalue && n_fn && h && g &&
              A && b && lda > 0,
          "invalid scatter arguments");
  const int rc = scatter_tables<double>(n_el_nodes, rows, cols, ctype, cvalue, n_fn, reversion, h,
                                        g, A, lda, b, false);
  if (rc) set_error("scatter: ctype must be 0, 1 or 10 with the matching columns present");
  return rc;
}

extern "C" int mfb_scatter_bem_harela_equation(int n_el_nodes, const int* rows, const int* cols,
                                               const int* ctype, const double* cvalue,
                                               const double* n_fn, int reversion, const double* h,
                                               const double* g, const double* u_inc,
                                               const double* t_inc, double* A, long long lda,
                                               double* b) {
  MFB_ARG(n_el_nodes > 0 && n_el_nodes <= 9 && rows && cols && ctype && cvalue && n_fn && h && g &&
              A && b && lda > 0 && (!u_inc) == (!t_inc),
          "invalid scatter arguments");
  typedef std::complex<double> z;
  const int rc = scatter_tables<z>(n_el_nodes, rows, cols, ctype, (const z*)cvalue, n_fn, reversion,
                                   (const z*)h, (const z*)g, (z*)A, lda, (z*)b, true);
  if (rc) {
    set_error("scatter: ctype must be 0, 1, 2, 3 or 10 with the matching columns present");
    return rc;
  }
  // incident wave field of the region (src/assemble_bem_harela_equation.f90:651-666):
  // b += h u_inc - g t_inc, u_inc / t_inc (3 x n_el_nodes) at the element's nodes
  if (u_inc) {
    const int ne = n_el_nodes;
    const z *H = (const z*)h, *G = (const z*)g, *U = (const z*)u_inc, *T = (const z*)t_inc;
    z* B = (z*)b;
    for (int il = 0; il < 3; il++)
      for (int ik = 0; ik < 3; ik++)
        for (int kn = 0; kn < ne; kn++) {
          const int o = kn + ne * il + 3 * ne * ik;
          B[rows[il]] += H[o] * U[ik + 3 * kn] - G[o] * T[ik + 3 * kn];
        }
  }
  return MFB_OK;
}

// Kernel blocks of independent (element, collocation point) pairs through the same kernels as the
// assembly: the GPU counterpart of fbem_bem_staela3d_sbie_auto / fbem_bem_harela3d_sbie_auto
// (bem_staela3d.f90:524-589, bem_harela3d.f90:1474-1538) for a batch of pairs.
static int pair_blocks(int npairs, const double* X, const double* xc, double qsi_re, int ns,
                       int n_ps, const int* ps_gln, int* classes, double* h, double* g,
                       bool harmonic, const KernelParams& kp) {
  MFB_ARG(npairs > 0 && X && xc && ps_gln && h && g, "null argument");
  // one element per pair (its own nodes) and one collocation point per pair; only the designated
  // pairs (element p, point p) are classified and integrated
  const int nn = 9 * npairs, ne = npairs;
  std::vector<double> x(3LL * nn);
  std::vector<int> conn(9LL * ne), ctype(3LL * nn, 1);
  std::vector<double> cval(3LL * nn, 0.0);
  CollocSet cs;
  cs.ncp = npairs;
  for (int p = 0; p < npairs; p++) {
    for (int k = 0; k < 9; k++) {
      conn[9 * p + k] = 9 * p + k;
      for (int c = 0; c < 3; c++) x[3 * (9 * p + k) + c] = X[27LL * p + 3 * k + c];
    }
    for (int c = 0; c < 3; c++) cs.xc.push_back(xc[3LL * p + c]);
    cs.cxi.push_back(0.0);
    cs.cxi.push_back(0.0);
    cs.crow.push_back(0);
    cs.cel.push_back(-1);
    cs.cshared.push_back(0);
  }
  mfb_mesh_t m = nullptr;
  int rc = mesh_create_impl(nn, ne, x.data(), conn.data(), ctype.data(), cval.data(), qsi_re, ns,
                            n_ps, ps_gln, 1e-6, &cs, 0.0, false, &m);
  if (rc) return rc;
  std::vector<int> pids(npairs);
  for (int p = 0; p < npairs; p++) pids[p] = p * npairs + p;
  std::vector<unsigned char> cls(npairs);
  int* dp = nullptr;
  unsigned char* dc = nullptr;
  double* dblk = nullptr;
  const int w = harmonic ? 2 : 1;
  auto done = [&](int code) {
    (void)hipFree(dp);
    (void)hipFree(dc);
    (void)hipFree(dblk);
    mfb_mesh_destroy(m);
    return code;
  };
  hipError_t he = hipMalloc(&dp, sizeof(int) * npairs);
  if (he == hipSuccess) he = hipMalloc(&dc, npairs);
  if (he == hipSuccess) he = hipMemcpy(dp, pids.data(), sizeof(int) * npairs, hipMemcpyHostToDevice);
  if (he == hipSuccess) {
    hipLaunchKernelGGL(k_classify_listed, dim3((npairs + 127) / 128), dim3(128), 0, 0, m->dm,
                       m->d_qsp, dp, npairs, dc);
    he = hipGetLastError();
  }
  if (he == hipSuccess) he = hipMemcpy(cls.data(), dc, npairs, hipMemcpyDeviceToHost);
  if (he != hipSuccess) {
    set_error("pair_blocks: %s", hipGetErrorString(he));
    return done(MFB_ERR_HIP);
  }
  std::vector<std::vector<int>> lst(m->n_cls);
  for (int p = 0; p < npairs; p++) {
    classes[p] = cls[p];
    lst[cls[p]].push_back(pids[p]);
  }
  std::vector<int> flat;
  for (int c = 0; c < m->n_cls; c++) {
    m->cls_off[c] = (long long)flat.size();
    m->cls_cnt[c] = (long long)lst[c].size();
    flat.insert(flat.end(), lst[c].begin(), lst[c].end());
  }
  m->cls_off[m->n_cls] = (long long)flat.size();
  he = hipMalloc(&m->lists, sizeof(int) * flat.size());
  if (he == hipSuccess)
    he = hipMemcpy(m->lists, flat.data(), sizeof(int) * flat.size(), hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMalloc((void**)&dblk, sizeof(double) * 162 * w * npairs);
  if (he != hipSuccess) {
    set_error("pair_blocks: %s", hipGetErrorString(he));
    return done(MFB_ERR_HIP);
  }
  rc = harmonic ? assemble<true>(m, kp, 0.0, nullptr, 0, nullptr, nullptr, dblk)
                : assemble<false>(m, kp, 0.0, nullptr, 0, nullptr, nullptr, dblk);
  if (rc == MFB_OK) {
    std::vector<double> hb(162LL * w * npairs);
    he = hipMemcpy(hb.data(), dblk, sizeof(double) * hb.size(), hipMemcpyDeviceToHost);
    if (he != hipSuccess) {
      set_error("pair_blocks: %s", hipGetErrorString(he));
      rc = MFB_ERR_HIP;
    }
    // blocks are in list order: slot q holds the pair of element flat[q] / npairs
    for (size_t q = 0; q < flat.size() && rc == MFB_OK; q++) {
      const long long p = flat[q] / npairs;
      for (int o = 0; o < 81 * w; o++) {
        h[81LL * w * p + o] = hb[162LL * w * q + o];
        g[81LL * w * p + o] = hb[162LL * w * q + 81 * w + o];
      }
    }
  }
  return done(rc);
}

extern "C" int mfb_pair_blocks_staela(int npairs, const double* X, const double* x_i, double mu,
                                      double nu, double qsi_relative_error, int qsi_ns_max,
                                      int n_precalsets, const int* precalset_gln, int* classes,
                                      double* h, double* g) {
  KernelParams kp;
  std::memset(&kp, 0, sizeof kp);
  kp.cteu1 = 1.0 / (16.0 * kPi * mu * (1.0 - nu));
  kp.cteu2 = 3.0 - 4.0 * nu;
  kp.ctet1 = -1.0 / (8.0 * kPi * (1.0 - nu));
  kp.ctet2 = 1.0 - 2.0 * nu;
  return pair_blocks(npairs, X, x_i, qsi_relative_error, qsi_ns_max, n_precalsets, precalset_gln,
                     classes, h, g, false, kp);
}

extern "C" int mfb_pair_blocks_harela(int npairs, const double* X, const double* x_i, double omega,
                                      double lambda_re, double lambda_im, double mu_re,
                                      double mu_im, double rho, double qsi_relative_error,
                                      int qsi_ns_max, int n_precalsets, const int* precalset_gln,
                                      int* classes, double* h, double* g) {
  KernelParams kp;
  std::memset(&kp, 0, sizeof kp);
  kp.harmonic = 1;
  host_har_params(lambda_re, lambda_im, mu_re, mu_im, rho, omega, &kp.h);
  return pair_blocks(npairs, X, x_i, qsi_relative_error, qsi_ns_max, n_precalsets, precalset_gln,
                     classes, h, g, true, kp);
}

// ------------------------------------------------------------------------------------------------
// interior points (src/calculate_internal_points_mechanics_static.f90,
// src/calculate_internal_points_mechanics_bem_harela.f90)
// ------------------------------------------------------------------------------------------------
struct mfb_points_s {
  mfb_mesh_t pm = nullptr;  // the boundary's elements with the points as collocation points
  int npts = 0;
};

// A point that coincides with a boundary node (distance below 1e-12 of the characteristic length
// of the node's first element) takes the node's displacement, as the reference does
// (src/calculate_internal_points_mechanics_bem_harela.f90:197-210).
template <typename T>
__global__ void k_snap_nodes(DevMesh m, const T* ub, T* out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= m.ncp) return;
  const double* xp = m.xc + 3 * p;
  for (int i = 0; i < m.nn; i++) {
    if (m.inc_ptr[i + 1] == m.inc_ptr[i]) continue;
    const double cl = m.elem[8 * m.inc_e[m.inc_ptr[i]]];
    const double dx = xp[0] - m.x[3 * i], dy = xp[1] - m.x[3 * i + 1], dz = xp[2] - m.x[3 * i + 2];
    if (sqrt(dx * dx + dy * dy + dz * dz) / cl < 1e-12)
      for (int k = 0; k < 3; k++) out[3 * p + k] = ub[3 * i + k];
  }
}

extern "C" int mfb_points_create(mfb_mesh_t mesh, int npts, const double* x, mfb_points_t* out) {
  MFB_ARG(mesh && npts > 0 && x && out, "bad argument");
  CollocSet cs;
  cs.ncp = npts;
  cs.xc.assign(x, x + 3LL * npts);
  cs.cxi.assign(2LL * npts, 0.0);
  cs.crow.assign(npts, 0);
  cs.cel.assign(npts, -1);
  cs.cshared.assign(npts, 0);
  const DevMesh& d = mesh->dm;
  mfb_mesh_t pm = nullptr;
  int rc = mesh_create_impl(d.nn, d.ne, mesh->h_x.data(), mesh->h_conn.data(),
                            mesh->h_ctype.data(), mesh->h_cval.data(), d.qsi_re, d.ns_max, d.n_ps,
                            mesh->h_ps.data(), mesh->h_tol, &cs, 0.0, true, &pm);
  if (rc) return rc;
  mfb_points_s* p = new mfb_points_s();
  p->pm = pm;
  p->npts = npts;
  *out = p;
  return MFB_OK;
}

extern "C" int mfb_points_destroy(mfb_points_t p) {
  if (!p) return MFB_OK;
  mfb_mesh_destroy(p->pm);
  delete p;
  return MFB_OK;
}

extern "C" int mfb_points_stats(mfb_points_t p, mfb_assembly_stats* st) {
  MFB_ARG(p, "null argument");
  return mfb_mesh_stats(p->pm, st);
}

extern "C" int mfb_points_pair_classes(mfb_points_t p, unsigned char* classes) {
  MFB_ARG(p, "null argument");
  return mfb_mesh_pair_classes(p->pm, classes);
}

template <bool H>
static int points_eval(mfb_points_t p, const KernelParams& kp, const double* ub, const double* tb,
                       double* out, void* stream) {
  typedef typename Val<H>::T T;
  MFB_ARG(p && ub && tb && out, "null argument");
  VecOut<T> v{(const T*)ub, (const T*)tb, (T*)out};
  int rc = assemble<H>(p->pm, kp, 0.0, nullptr, 0, nullptr, stream, nullptr, 0.0, v);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_snap_nodes<T>, dim3((p->npts + 127) / 128), dim3(128), 0, s, p->pm->dm,
                     (const T*)ub, (T*)out);
  MFB_HIP(hipGetLastError());
  MFB_HIP(hipStreamSynchronize(s));
  return MFB_OK;
}

extern "C" int mfb_points_displacement_staela(mfb_points_t p, double mu, double nu,
                                              const double* u_bnd, const double* t_bnd,
                                              double* u_out, void* stream) {
  MFB_ARG(mu > 0 && nu < 0.5 && nu > -1.0, "invalid elastic constants");
  return points_eval<false>(p, static_params(mu, nu), u_bnd, t_bnd, u_out, stream);
}

extern "C" int mfb_points_displacement_harela(mfb_points_t p, double omega, double lambda_re,
                                              double lambda_im, double mu_re, double mu_im,
                                              double rho, const double* u_bnd,
                                              const double* t_bnd, double* u_out, void* stream) {
  MFB_ARG(omega > 0 && rho > 0 && mu_re > 0, "invalid harmonic parameters (omega > 0 required)");
  KernelParams kp;
  std::memset(&kp, 0, sizeof kp);
  kp.harmonic = 1;
  host_har_params(lambda_re, lambda_im, mu_re, mu_im, rho, omega, &kp.h);
  return points_eval<true>(p, kp, u_bnd, t_bnd, u_out, stream);
}
