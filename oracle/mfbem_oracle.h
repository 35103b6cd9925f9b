/* CPU restatement of MultiFEBE's 3D BEM influence-matrix assembly (static Kelvin and time-harmonic
 * visco-elastic Stokes kernels, nodal SBIE on quad9 meshes).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle: tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, as the checker and as the timed CPU baseline.  The
 * product (multifebe_amd, libmfbem.so) never links, loads or calls it.
 *
 * Array conventions (all 0-based, Fortran column-major kept where the reference has an array):
 *   element node coordinates  x[3*k + c]            == x(c,k)         (3 x 9)
 *   kernel blocks             h[n + 9*l + 27*k]     == h(n,l,k)       (9 x 3 x 3)
 *   dense matrix              A[row + n_dof*col]    == A(row,col)
 *   free term                 C[l + 3*k]            == c(l,k)
 */
#ifndef MFBEM_ORACLE_H
#define MFBEM_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Quadrature tables as the reference holds them (rule n = 1..32, 32 entries per rule). */
void orc_gl_tables(double *gl11_xi, double *gl11_w, double *gl01_xi, double *gl01_w);

/* E_0..E_6 of exp(z) (numerical.f90:1258-1330). E is 7 complex values (re,im interleaved). */
void orc_zexp(double zr, double zi, double *E);

/* Fixed rule: Gauss-Legendre order gln on each of the 2^s x 2^s uniform sub-quads (no Telles). */
int orc_pair_rule_sta(const double *x, const double *x_i, double mu, double nu, int reverse,
                      int gln, int s, double *h, double *g);
int orc_pair_rule_har(const double *x, const double *x_i, double lam_r, double lam_i, double mu_r,
                      double mu_i, double rho, double omega, int reverse, int gln, int s,
                      double *h, double *g);

/* Mantic free term (bem_staela3d.f90:228-405); n,t are 3 x m. */
void orc_freeterm(int m, const double *n, const double *t, double tol, double nu, double *C);

/* Element data (src/build_data_of_be_elements.f90:70-112): characteristic length, gln_far, ball. */
void orc_element_data(const double *x, double qsi_re, double *csize, int *n_phi, double *centre,
                      double *radius);

/* Nearest element point (geometry.f90:5519-5572) for a quad9. */
void orc_nearest_quad9(const double *x, double cl, const double *x_i, double *barxi, double *rmin,
                       double *d, int *method);

/* One (collocation point, element) pair through the reference's automatic integration.
 * trace[0]=delta (1 singular), trace[1]=gln_near, trace[2]=gln, trace[3]=ps (1-based, 0 = adaptive)
 * dtrace = barxi(2), d, adaptive leaves.  h,g are 9x3x3 (static: real; harmonic: complex interleaved, 2*81 doubles). */
int orc_pair_sta(const double *x, const double *x_i, double mu, double nu, double qsi_re, int ns,
                 int reverse, int n_ps, const int *ps_gln, int *trace, double *dtrace, double *h,
                 double *g);
int orc_pair_har(const double *x, const double *x_i, double lam_r, double lam_i, double mu_r,
                 double mu_i, double rho, double omega, double qsi_re, int ns, int reverse, int n_ps,
                 const int *ps_gln, int *trace, double *dtrace, double *h, double *g);

/* Whole-boundary assembly of A x = b for one closed quad9 boundary with nodal collocation.
 * conn: 9 x ne, 0-based node ids.  ctype: 3 x nn (0 u known, 1 t known, 10 normal pressure),
 * cval: 3 x nn prescribed values.  Unknown / equation of node i, direction k at index 3*i+k.
 * Only equations of collocation nodes in [node_begin, node_end) are assembled (the whole system
 * when node_begin=0, node_end=nn); A must be zeroed by the caller.
 * stats (may be NULL): [0] pairs, [1] singular pairs, [2] precalculated-rule pairs,
 *                      [3] adaptive pairs, [4] quadrature points, [5] adaptive leaves.
 * Harmonic A, b are complex interleaved. */
int orc_assemble_sta(int nn, int ne, const double *xn, const int *conn, const int *ctype,
                     const double *cval, double mu, double nu, double qsi_re, int ns, int n_ps,
                     const int *ps_gln, double tol, int node_begin, int node_end, double *A,
                     double *b, long long *stats);
int orc_assemble_har(int nn, int ne, const double *xn, const int *conn, const int *ctype,
                     const double *cval, double lam_r, double lam_i, double mu_r, double mu_i,
                     double rho, double omega, double nu, double qsi_re, int ns, int n_ps,
                     const int *ps_gln, double tol, int node_begin, int node_end, double *A,
                     double *b, long long *stats);

/* Integration class of every (element e, node i) pair, ne*nn bytes at e*nn + i: 0..n_ps-1 the
 * precalculated dataset, n_ps adaptive, n_ps+1 singular. */
int orc_collocation(int nn, int ne, const int *conn, const double *xn, double delta, double *xc,
                    int *crow, int *cel, double *cxi, unsigned char *is_mca);
void orc_pair_classes_pts(int npt, const double *xp, int ne, const double *xn, const int *conn,
                          double qsi_re, int n_ps, const int *ps_gln, unsigned char *out);
int orc_points_sta(int nn, int ne, const double *xn, const int *conn, int npt, const double *xp,
                   double mu, double nu, double qsi_re, int ns, int n_ps, const int *ps_gln,
                   const double *ub, const double *tb, double *out);
int orc_points_har(int nn, int ne, const double *xn, const int *conn, int npt, const double *xp,
                   double lam_r, double lam_i, double mu_r, double mu_i, double rho, double omega,
                   double qsi_re, int ns, int n_ps, const int *ps_gln, const double *ub,
                   const double *tb, double *out);
void orc_pair_classes(int nn, int ne, const double *xn, const int *conn, double qsi_re, int n_ps,
                      const int *ps_gln, unsigned char *out);

#ifdef __cplusplus
}
#endif
#endif
