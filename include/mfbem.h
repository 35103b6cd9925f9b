/* libmfbem: MI355X (gfx950) BEM influence-matrix assembly and dense solve for 3D elastostatic
 * (Kelvin) and time-harmonic visco-elastic (complex Stokes) regions.
 *
 * C ABI: plain pointers and sizes only.  Every entry point returns 0 on success or a negative
 * MFB_ERR_* code; mfb_last_error() then describes the failure (the reference prints and stops,
 * lib/fbem/src/string_handling.f90:289-333, so a Fortran shim maps a nonzero return to
 * fbem_error_message).  Entry points are not re-entrant; one host thread drives a mesh handle.
 *
 * Layouts
 *   node coordinates  x[3*i + c]                      (nn x 3, row-major)
 *   connectivity      conn[9*e + k], 0-based          (ne x 9, quad9 local order of the reference)
 *   conditions        ctype[3*i + k] in {0 u known, 1 t known, 10 normal pressure}, cval[3*i + k]
 *   unknown / equation of node i, direction k         index 3*i + k
 *   dense matrices    column-major, A[row + lda*col]; complex values interleaved (re, im), which is
 *                     the memory layout of Fortran complex(8)
 *   "device" pointers are HIP device allocations on the current device; "host" pointers are
 *   ordinary host memory.
 */
#ifndef MFBEM_H
#define MFBEM_H

#ifdef __cplusplus
extern "C" {
#endif

#define MFB_OK 0
#define MFB_ERR_ARG (-1)        /* invalid argument                                        */
#define MFB_ERR_HIP (-2)        /* HIP runtime error                                       */
#define MFB_ERR_SINGULAR (-3)   /* zero pivot in the LU (LAPACK info > 0)                  */
#define MFB_ERR_UNSUPPORTED (-4)/* option not provided by this build                        */
#define MFB_ERR_GEOMETRY (-5)   /* invalid free-term normal/tangent set, degenerate element */
#define MFB_ERR_LIMIT (-6)      /* adaptive subdivision produced more leaves than the cap   */

typedef struct mfb_mesh_s *mfb_mesh_t;

typedef struct {
  long long pairs;      /* collocation point x element pairs                               */
  long long singular;   /* pairs integrated by the polar (singular) rule                   */
  long long pre;        /* pairs integrated with a precalculated Gauss-Legendre dataset     */
  long long adaptive;   /* pairs integrated by Telles + adaptive subdivision               */
  long long points;     /* quadrature points evaluated by the last assembly                */
  long long leaves;     /* adaptive sub-elements integrated by the last assembly           */
  long long decisions;  /* class histogram checksum: sum over pairs of (class+1)*(pair%97+1) */
  long long regular_points; /* quadrature points of the precalculated-rule pairs (per assembly) */
  long long learned;        /* 1 when the last assembly chose the rules with the network     */
  long long judge_pairs;    /* pairs the judge checked in the last (learned) assembly         */
  double judge_max_err;     /* judge: largest block error of the network's rules (relative)   */
  double judge_mean_err;    /* judge: mean block error                                        */
  long long judge_ref_pairs;/* judge pairs where the reference's estimator chose a
                               precalculated rule ...                                          */
  double judge_ref_max_err; /* ... and that rule's largest block error on them              */
  long long judge_promoted; /* rules the judge's action promoted in the last assembly (then
                               assembled once more with them)                                  */
  double judge_first_max_err; /* the judge's largest error before those promotions            */
} mfb_assembly_stats;

typedef struct {
  int info;             /* LAPACK-style: 0 ok, k > 0 zero pivot at column k               */
  int precision_mode;   /* mode that produced the returned solution                         */
  int refinement_steps; /* FP64 iterative-refinement sweeps (mixed mode)                    */
  int fell_back;        /* bit 1: the mixed mode stalled and the FP64 LU was used; bit 2:   */
                        /* a persistent kernel's bounded wait expired and the work was redone */
                        /* with one launch per step (column panels, block solves)            */
  double backward_error;/* max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf) */
  double factor_ms;     /* time of the factorisation(s)                                      */
  double solve_ms;      /* time of the triangular solves and refinement                     */
  double rcond;         /* solve_lse condition stage: ?gecon's reciprocal 1-norm condition   */
  double ferr;          /* solve_lse refine stage: ?gerfs' largest forward error bound       */
} mfb_solve_report;

#define MFB_PREC_FP64 0          /* FP64 LU (dgesv / zgesv semantics)                        */
#define MFB_PREC_FP32_REFINE 1   /* FP32 LU on MFMA + FP64 iterative refinement, FP64 fallback */
/* FP32 LU whose trailing updates run as one FP16 GEMM of three hi/lo products with FP32
 * accumulation (about FP32 accuracy, ~2^-21 |L| |U|) + FP64 iterative refinement, FP64 fallback */
#define MFB_PREC_SPLIT16_REFINE 2

const char *mfb_last_error(void);

/* Upload a quad9 boundary and precompute everything that depends on geometry only.  Collocation
 * follows the reference's default formulation (src/assign_default_bem_formulation.f90:74-93):
 * nodal SBIE on every node of a closed surface; a node on an open edge (an element edge no other
 * element shares) takes SBIE-MCA -- one collocation point in every element holding it, at
 * xi_node (1 - 0.05), all adding to the node's equations, free term (1/2) phi(xi) -- so truncated
 * open surfaces (a half-space free surface) are assembled as the reference assembles them.
 * Precomputed:
 * element characteristic lengths, gln_far and bounding balls (src/build_data_of_be_elements.f90:70-112),
 * precalculated Gauss-Legendre datasets (lib/fbem/src/bem_general.f90:163-759), node normals, and
 * the integration class of every (collocation node, element) pair (bem_staela3d.f90:549-587).
 * qsi_relative_error, qsi_ns_max, precalset_gln and geometric_tolerance are the reference's
 * [settings] (src/read_settings.f90:77-217).  Replaces the per-element setup that
 * src/build_lse_mechanics_bem_staela.f90:659-751 repeats for every element and frequency. */
int mfb_mesh_create(int nn, int ne, const double *x, const int *conn, const int *ctype,
                    const double *cval, double qsi_relative_error, int qsi_ns_max,
                    int n_precalsets, const int *precalset_gln, double geometric_tolerance,
                    mfb_mesh_t *mesh);
int mfb_mesh_destroy(mfb_mesh_t mesh);
int mfb_mesh_stats(mfb_mesh_t mesh, mfb_assembly_stats *stats);
/* Durations (ms, HIP events on the assembly stream) of the kernel groups of the last assembly:
 * ms[0] precalculated-rule pairs, ms[1] adaptive pairs, ms[2] singular pairs, ms[3] zeroing and
 * free terms; regular_launches: kernel launches behind ms[0].  No counterpart in the reference
 * (instrumentation of this build). */
int mfb_mesh_timings(mfb_mesh_t mesh, double *ms, int *regular_launches);
/* integration class of every (collocation point, element) pair, host array of ncp*ne bytes,
 * index e*ncp + cp: 0..n_ps-1 precalculated dataset, n_ps adaptive, n_ps+1 singular.  (ncp = nn
 * on a closed surface, point cp = node cp.)  Pair ids taken or returned by the functions below are
 * e*ncp + cp as well. */
int mfb_mesh_pair_classes(mfb_mesh_t mesh, unsigned char *classes);
/* the same layout for the classes the quadrature network chose in the last assembly of a mesh
 * with a network (mfb_mesh_set_qnet), after the judge's promotions.  No counterpart in the
 * reference (the network replaces its estimator). */
int mfb_mesh_learned_classes(mfb_mesh_t mesh, unsigned char *classes);
/* Collocation points (src/build_data_at_collocation_points.f90): their number, and (host arrays,
 * each may be null) coordinates xc (ncp x 3), the node whose equations each adds to, and the
 * element holding an MCA point (-1 for a nodal point).  No counterpart entry in the reference
 * (its element(se)%x_i_sbie / x_i_sbie_mca arrays). */
int mfb_mesh_collocation(mfb_mesh_t mesh, int *ncp, double *xc, int *crow, int *cel);
/* Node normals n_fn (host, nn x 3): the normalised sum of the unit normals of the elements at the
 * node, the direction of a ctype-10 pressure (node%n_fn of the reference). */
int mfb_mesh_node_normals(mfb_mesh_t mesh, double *nfn);

/* Assemble the whole static system A x = b into DEVICE memory (A: n_dof x n_dof, lda >= n_dof,
 * b: n_dof).  Replaces build_lse_mechanics_bem_staela (src/build_lse_mechanics_bem_staela.f90:22-655)
 * including every call of assemble_bem_staela_equation (src/assemble_bem_staela_equation.f90:22-329,
 * BE ordinary boundary) and the free terms.  A and b are overwritten (not accumulated).
 * stream: hipStream_t or NULL for the default stream. */
int mfb_assemble_staela(mfb_mesh_t mesh, double mu, double nu, double *A, long long lda, double *b,
                        void *stream);

/* Harmonic counterpart (src/build_lse_mechanics_bem_harela.f90:22-970 with
 * src/assemble_bem_harela_equation.f90:78-113): complex A, b interleaved.  lambda, mu are the
 * complex Lame constants (mu* = mu (1 + 2 i xi), src/read_regions.f90:717-720), nu the real
 * Poisson ratio used by the free term. */
int mfb_assemble_harela(mfb_mesh_t mesh, double omega, double lambda_re, double lambda_im,
                        double mu_re, double mu_im, double rho, double nu, double *A,
                        long long lda, double *b, void *stream);

/* Batches of whole-boundary assemblies (no counterpart in the reference, whose driver assembles
 * one region at a time: the batched configurations of many independent problems and frequencies).
 * A job is one mesh at one material / frequency into its own A (lda), b (device).  harmonic = 0:
 * static with mu_re (shear modulus) and nu; 1: harmonic as mfb_assemble_harela.  The reference's
 * rule choice only (a mesh with a network is rejected). */
typedef struct {
  mfb_mesh_t mesh;
  int harmonic;
  double omega, lambda_re, lambda_im, mu_re, mu_im, rho, nu;
  double *A;
  long long lda;
  double *b;
} mfb_assembly_job;
/* every job's kernels on `stream` back to back, with no host synchronisation between them; one
 * synchronisation at the end checks the error flags. */
int mfb_assemble_batch(int njobs, const mfb_assembly_job *jobs, void *stream);
/* The same sequence captured once into a HIP graph (memcpy + kernel nodes only) and replayed:
 * set_params rewrites the materials / frequencies of the same jobs (same meshes, buffers and
 * kinds) for the next launch; check synchronises and reports the kernels' error flags.  Jobs of
 * one graph run in order; a graph must not be launched twice at once. */
typedef struct mfb_batch_graph_s *mfb_batch_graph_t;
int mfb_batch_graph_create(int njobs, const mfb_assembly_job *jobs, mfb_batch_graph_t *graph);
int mfb_batch_graph_set_params(mfb_batch_graph_t graph, int njobs, const mfb_assembly_job *jobs);
int mfb_batch_graph_launch(mfb_batch_graph_t graph, void *stream);
int mfb_batch_graph_check(mfb_batch_graph_t graph);
int mfb_batch_graph_destroy(mfb_batch_graph_t graph);

/* Kernel blocks h, g (column-major 9x3x3 each, complex interleaved for harela) of npairs
 * independent (element, collocation point) pairs, computed by the same GPU kernels as the
 * assembly: the batched counterpart of fbem_bem_staela3d_sbie_auto (lib/fbem/src/bem_staela3d.f90:524-589)
 * and fbem_bem_harela3d_sbie_auto (lib/fbem/src/bem_harela3d.f90:1474-1538) with reverse = false.
 * X: npairs x 27 element node coordinates, x_i: npairs x 3; classes (host, npairs) receives the
 * integration class (0..n_ps-1 precalculated dataset, n_ps adaptive, n_ps+1 singular).  Host arrays. */
int mfb_pair_blocks_staela(int npairs, const double *X, const double *x_i, double mu, double nu,
                           double qsi_relative_error, int qsi_ns_max, int n_precalsets,
                           const int *precalset_gln, int *classes, double *h, double *g);
int mfb_pair_blocks_harela(int npairs, const double *X, const double *x_i, double omega,
                           double lambda_re, double lambda_im, double mu_re, double mu_im,
                           double rho, double qsi_relative_error, int qsi_ns_max, int n_precalsets,
                           const int *precalset_gln, int *classes, double *h, double *g);

/* Per-pair scatter of one 9x3x3 kernel block into HOST A, b: the body of
 * assemble_bem_staela_equation for an ordinary BE boundary (src/assemble_bem_staela_equation.f90:69-98).
 * h, g are column-major (9,3,3); node_col is the collocation node; enodes the 9 element nodes;
 * node_normal (nn x 3) is used by ctype 10.  reversion: sb_int_reversion. */
int mfb_assemble_bem_staela_equation(int nn, const int *ctype, const double *cval,
                                     const double *node_normal, int reversion, int node_col,
                                     const int *enodes, const double *h, const double *g, double *A,
                                     long long lda, double *b);
/* complex version (src/assemble_bem_harela_equation.f90:78-113); h, g, A, b complex interleaved,
 * cval real */
int mfb_assemble_bem_harela_equation(int nn, const int *ctype, const double *cval,
                                     const double *node_normal, int reversion, int node_col,
                                     const int *enodes, const double *h, const double *g, double *A,
                                     long long lda, double *b);

/* The same scatter with the reference's own DOF tables instead of the 3 i + k numbering (the body
 * of assemble_bem_staela_equation / assemble_bem_harela_equation for an ordinary BE boundary,
 * src/assemble_bem_staela_equation.f90:69-98, src/assemble_bem_harela_equation.f90:78-113):
 * n_el_nodes element nodes; rows[3] = node(sn_col)%row(1:3, eq_index) - 1; cols[6 kn + k] =
 * node(sn)%col(k + 1, 1) - 1 of element node kn (k < 3: u_k, k >= 3: t_k; -1 where absent);
 * ctype, cvalue (real, or complex interleaved for harela), n_fn: (n_el_nodes x 3) of the element's
 * nodes; h, g: (n_el_nodes, 3, 3) column-major; A (lda), b: HOST arrays, column-major, 0-based.
 * harela also takes ctype 2 / 3 (local-axes conditions: u and t both unknown). */
int mfb_scatter_bem_staela_equation(int n_el_nodes, const int *rows, const int *cols,
                                    const int *ctype, const double *cvalue, const double *n_fn,
                                    int reversion, const double *h, const double *g, double *A,
                                    long long lda, double *b);
/* harela also adds the region's incident wave field, b += h u_inc - g t_inc
 * (src/assemble_bem_harela_equation.f90:651-666), when u_inc, t_inc (3 x n_el_nodes complex at
 * the element's nodes) are given (both null: no incident field). */
int mfb_scatter_bem_harela_equation(int n_el_nodes, const int *rows, const int *cols,
                                    const int *ctype, const double *cvalue, const double *n_fn,
                                    int reversion, const double *h, const double *g,
                                    const double *u_inc, const double *t_inc, double *A,
                                    long long lda, double *b);

/* Dense solve on DEVICE memory.  A (n x n, lda) is overwritten by its LU factors (FP64 mode) or
 * left intact (mixed mode keeps the FP64 A for the residual); B (n x nrhs, ldb) by the solution;
 * ipiv (device, n ints) receives LAPACK-style 1-based pivots in FP64 mode.  The FP32 work copy
 * of the mixed mode is allocated inside the call. */
int mfb_dgesv_device(int n, int nrhs, double *A, long long lda, int *ipiv, double *B, long long ldb,
                     int precision_mode, mfb_solve_report *report, void *stream);
int mfb_zgesv_device(int n, int nrhs, double *A, long long lda, int *ipiv, double *B, long long ldb,
                     int precision_mode, mfb_solve_report *report, void *stream);

/* Drop-in for solve_lse_r / solve_lse_c (src/solve_lse_r.f90:25-46, src/solve_lse_c.f90:25-46) on
 * HOST arrays: the same argument list (logicals as int, character(1) as char).
 *   factorize = 1: A (n x n, column-major) is uploaded, optionally equilibrated (scaling = 1:
 *     LAPACK ?geequ + ?laqge, r, c and equed returned as the reference's ?geequ/?laqge return
 *     them, src/solve_lse_c.f90:81-117), factorised on the GPU and b (n x n_rhs) solved.  The
 *     factor stays on the device, keyed by the address of A, n and precision_mode.  A is
 *     overwritten by LAPACK-format LU factors and ipiv by 1-based pivots (FP64 mode: those of
 *     ?getrf; mixed mode: the FP32 factors of the scaled matrix written as FP64 factors of A).
 *   factorize = 0: b is solved with the kept factor of the same A (src/multifebe.f90:119) --
 *     only the triangular solves (and, in mixed mode, the FP64 refinement) run; with scaling the
 *     caller's r, c and equed are applied to b as the reference does.  MFB_ERR_ARG when no kept
 *     factor matches.
 * condition = 1 (with factorize = 1, src/solve_lse_c.f90:140-165): report->rcond = ?gecon's
 *   estimate (norm '1') from the FP64 factors and ||Ao||_1.
 * refine = 1 (src/solve_lse_c.f90:191-206): ?gerfs on the solutions -- FP64 residual with the
 *   unfactored matrix, corrections while LAPACK's rules allow (at most 5), backward error in
 *   report->backward_error, forward error bound in report->ferr (largest over the right-hand
 *   sides), corrections applied in report->refinement_steps.
 * With condition or refine, Aodim must be n_dof: factorize = 1 writes the (scaled) matrix into Ao
 *   as the reference does, and the factorisation is FP64 whatever precision_mode asks; a
 *   factorize = 0 refine needs such a factor (its Ao is kept on the device). */
int mfb_solve_lse_r(int n_dof, double *A, int *ipiv, int aodim, double *Ao, char *equed, double *r,
                    double *c, int n_rhs, double *b, int factorize, int scaling, int condition,
                    int refine, int precision_mode, mfb_solve_report *report);
int mfb_solve_lse_c(int n_dof, double *A, int *ipiv, int aodim, double *Ao, char *equed, double *r,
                    double *c, int n_rhs, double *b, int factorize, int scaling, int condition,
                    int refine, int precision_mode, mfb_solve_report *report);

/* Release everything the library keeps between calls (kept factors, library handles, the second
 * stream, device workspaces).  Meshes and networks are the caller's to destroy.  Later calls
 * re-create what they need.  No counterpart in the reference (process teardown of this build). */
int mfb_shutdown(void);

/* Schedule of the dense solves (mfb_dgesv_device / mfb_zgesv_device, mfb_solve_lse_*): the choices
 * the batched configuration's autotuner makes.  No counterpart in the reference (LAPACK has no
 * schedule).  lookahead: 1 the trailing update beyond the next panel on a second stream, 0 all on
 * the caller's stream; la_chunk: columns per second-stream trailing-update GEMM (rounded down to a
 * multiple of 256), 0 the built-in choice by type; trsv_parts: most workgroups per row block of
 * the one-right-hand-side triangular solves (1, 2 or 4; fewer where that grid is not co-resident).
 * -1 in a field: the environment's (MFB_LOOKAHEAD, MFB_LA_CHUNK, MFB_TRSV_SPLIT) or built-in
 * default.  Process-wide, read by every later call.  Every schedule gives LAPACK's pivots in FP64
 * mode and solutions that agree to rounding. */
int mfb_solver_set_schedule(int lookahead, int la_chunk, int trsv_parts);
int mfb_solver_get_schedule(int *lookahead, int *la_chunk, int *trsv_parts);

/* ---------------------------------------------------------------------------------------------
 * Interior points (src/calculate_internal_points_mechanics_static.f90:76-89,
 * src/calculate_internal_points_mechanics_bem_harela.f90:160-172, 355-363): the displacement of
 * the region at points inside it from the boundary solution,
 *   u_l(x) = sum over elements e, nodes n, directions k of g(n,l,k) t_k(n) - h(n,l,k) u_k(n),
 * each (point, element) pair integrated by the same decision sequence and kernels as the boundary
 * pairs (sbie_auto).  A point on a boundary node takes the node's displacement.
 * ------------------------------------------------------------------------------------------- */
typedef struct mfb_points_s *mfb_points_t;
/* x: HOST (npts x 3).  The point set uses the mesh's elements and settings (the mesh may be
 * destroyed afterwards); pairs are classified at creation. */
int mfb_points_create(mfb_mesh_t mesh, int npts, const double *x, mfb_points_t *points);
int mfb_points_destroy(mfb_points_t points);
int mfb_points_stats(mfb_points_t points, mfb_assembly_stats *stats);
/* integration class of every (point, element) pair, host array npts*ne bytes, index e*npts + p */
int mfb_points_pair_classes(mfb_points_t points, unsigned char *classes);
/* u_bnd, t_bnd: DEVICE (nn x 3) displacements and tractions at the boundary nodes (complex
 * interleaved for harela); u_out: DEVICE (npts x 3), overwritten. */
int mfb_points_displacement_staela(mfb_points_t points, double mu, double nu, const double *u_bnd,
                                   const double *t_bnd, double *u_out, void *stream);
int mfb_points_displacement_harela(mfb_points_t points, double omega, double lambda_re,
                                   double lambda_im, double mu_re, double mu_im, double rho,
                                   const double *u_bnd, const double *t_bnd, double *u_out,
                                   void *stream);

/* ---------------------------------------------------------------------------------------------
 * Learned quadrature-rule selection (no counterpart call in the reference: the network replaces
 * the fitted estimator fbem_qs_n_estimation_standard, lib/fbem/src/quasisingular_integration.f90:
 * 294-715, where sbie_auto chooses a pair's rule, lib/fbem/src/bem_staela3d.f90:549-587).
 * MLP: MFB_QNET_INPUTS features -> 3 x 64 tanh -> MFB_QNET_CLASSES logits, FP32 on the GPU.
 * Classes c = 0..7: precalculated Gauss-Legendre rule of order 2 + c; c = 8: adaptive rule.
 * X, y, probs and classes of the functions below are DEVICE arrays (row-major, n rows).
 * ------------------------------------------------------------------------------------------- */
#define MFB_QNET_INPUTS 5
#define MFB_QNET_CLASSES 9
typedef struct mfb_qnet_s *mfb_qnet_t;

int mfb_qnet_param_count(void);
int mfb_qnet_create(unsigned long long seed, mfb_qnet_t *qnet);
int mfb_qnet_destroy(mfb_qnet_t qnet);
/* host_params: mfb_qnet_param_count() floats; set resets the optimizer state */
int mfb_qnet_get_params(mfb_qnet_t qnet, float *host_params);
int mfb_qnet_set_params(mfb_qnet_t qnet, const float *host_params);
/* gradient of the mean softmax cross-entropy over the n rows (grad, loss: HOST outputs) */
int mfb_qnet_gradient(mfb_qnet_t qnet, const float *X, const int *y, int n, float *grad,
                      float *loss);
/* mfb_qnet_gradient without any device allocation, for callers that cannot allocate (code under
 * HIP graph capture, tight training or serving loops).  No counterpart in the reference.  The
 * per-workgroup partial sums live in a fixed device array of the library (9.5 MB per GPU, held
 * until the process ends); n is limited to 16384 rows (MFB_ERR_LIMIT beyond).  Calls that
 * overlap on one GPU share that array and race.  Same kernels and summation order as
 * mfb_qnet_gradient: the same grad and loss. */
int mfb_qnet_gradient_noalloc(mfb_qnet_t qnet, const float *X, const int *y, int n, float *grad,
                              float *loss);
/* Adam (lr, beta 0.9/0.999, eps 1e-8) over shuffled mini-batches; loss_per_epoch (HOST, epochs
 * floats, may be null) receives the mean mini-batch loss of each epoch */
int mfb_qnet_train(mfb_qnet_t qnet, const float *X, const int *y, int n, int epochs, int batch,
                   float lr, unsigned long long seed, float *loss_per_epoch);
/* probabilities (n x classes, may be null) and the chosen class per row (may be null): the
 * cheapest class whose cumulative probability reaches confidence */
int mfb_qnet_infer(mfb_qnet_t qnet, const float *X, int n, float confidence, float *probs,
                   int *classes, void *stream);

/* Rule choice with the network for the following assemblies of this mesh (qnet = NULL: back to
 * the reference's estimator).  Every assembly then classifies all pairs with the network: class
 * MFB_QNET_CLASSES - 1, or no class reaching the confidence, takes the reference's adaptive rule.
 * The judge integrates judge_samples of the network's pairs with the adaptive rule at
 * judge_qsi_relative_error and reports the largest error in mfb_mesh_stats.  The mesh does not
 * own qnet; keep it alive while it is set. */
int mfb_mesh_set_qnet(mfb_mesh_t mesh, mfb_qnet_t qnet, float confidence,
                      double judge_qsi_relative_error, int judge_samples);
/* The judge's action (no counterpart in the reference, whose estimator has no judge): with
 * promote = 1, a precalculated rule whose judged pairs miss the mesh's qsi_relative_error is
 * replaced -- for the network classes that chose it, in this and every later assembly of the
 * mesh -- by the cheapest longer rule that meets the tolerance on all of them, or by the adaptive
 * rule; the assembly is then run once more with the new choice, and mfb_mesh_stats reports the
 * promotions and the judge's error before and after.  0 (default): the judge only reports. */
int mfb_mesh_set_judge_action(mfb_mesh_t mesh, int promote);

/* Features of listed pairs of a mesh (pair id = element * nn + node, DEVICE int array) for the
 * network: log10(d)/2 (d as sbie_auto defines it), log2(aspect), curvature (|d2x/dxi2| / h),
 * k h / 4 (wavenumber k: 0 for the static kernel), edge proximity of the nearest point.
 * feats: DEVICE, np x MFB_QNET_INPUTS. */
int mfb_mesh_pair_features(mfb_mesh_t mesh, const int *pairs, int np, double wavenumber,
                           float *feats, void *stream);
/* Labels for the network: every precalculated rule of the mesh is compared with the reference's
 * adaptive rule run at the mesh's own qsi_relative_error (create the labelling mesh with a tight
 * one); errors[q * n_precalsets + r] = max(|h_r - h| / max|h|, |g_r - g| / max|g|) and labels[q]
 * = class of the cheapest rule whose error is <= tau (MFB_QNET_CLASSES - 1 when none is), -1
 * for singular pairs.  labels, errors (may be null): DEVICE arrays. */
int mfb_mesh_pair_labels_staela(mfb_mesh_t mesh, double mu, double nu, const int *pairs, int np,
                                double tau, int *labels, double *errors, void *stream);
int mfb_mesh_pair_labels_harela(mfb_mesh_t mesh, double omega, double lambda_re, double lambda_im,
                                double mu_re, double mu_im, double rho, const int *pairs, int np,
                                double tau, int *labels, double *errors, void *stream);

#ifdef __cplusplus
}
#endif
#endif
