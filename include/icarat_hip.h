/* icarat_hip.h -- C ABI of the MI355X (gfx950) implementation of ICARAT's
 * linear-elastic SIMP compliance loop on a voxel hexa8 mesh.
 *
 * Conventions
 *  - Every entry point returns an int status (ICARAT_HIP_OK == 0).  On a
 *    non-zero status the text of the error is kept in the context and
 *    icarat_hip_last_error() returns it.  No entry point aborts the process.
 *  - Host pointers are borrowed for the duration of the call only.  Device
 *    memory is owned by the context.
 *  - Vectors on DOFs are in the reference's "numeq" order (free DOFs numbered
 *    in node order, reference io/boundary_condition_base.hpp:35-95) and
 *    element vectors in the reference's element order (x fastest, reference
 *    io/mesh_icarat.hpp:480-511).
 *  - A context is single-threaded: calls on one context must not overlap.
 *    Distinct contexts are independent.
 */
#ifndef ICARAT_HIP_H
#define ICARAT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ICARAT_HIP_OK = 0,
  ICARAT_HIP_INVALID_ARGUMENT = 1,
  ICARAT_HIP_HIP_ERROR = 2,
  ICARAT_HIP_NOT_READY = 3,
  ICARAT_HIP_OUT_OF_MEMORY = 4,
  ICARAT_HIP_NOT_CONVERGED = 5,
  ICARAT_HIP_NUMERICAL = 6,
  ICARAT_HIP_INTERNAL = 7
};

/* Precision of the element contraction used inside a solve. */
enum {
  ICARAT_HIP_FP64 = 0,         /* plain FP64 throughout */
  ICARAT_HIP_FP32 = 1,         /* FP32 MFMA contraction inside FP64 refinement */
  ICARAT_HIP_BF16 = 2,         /* BF16 MFMA contraction (FP32 accumulate) inside FP64 refinement */
  ICARAT_HIP_FP64_MFMA = 3     /* icarat_hip_apply_K only: FP64 MFMA element contraction */
};

/* Preconditioner of the (inner) Krylov solve. */
enum {
  ICARAT_HIP_PRECOND_JACOBI = 0,   /* reference "Scaling_CG" (analysis/solver/CG.cpp:56-100) */
  ICARAT_HIP_PRECOND_LEARNED = 1,  /* U-Net correction in flexible PCG, Jacobi fallback */
  ICARAT_HIP_PRECOND_MG = 2        /* geometric multigrid V-cycle (see icarat_hip_mg_setup) */
};

typedef struct icarat_hip_ctx icarat_hip_ctx;

/* Box mesh + material.  Replaces MeshIcarat::setParameter/meshBox for hexa8
 * (io/mesh_icarat.hpp:72-139, 459-513) and Elastic::setParameter
 * (problem/elasticity.cpp:19-33): young1/young2 are the reference's
 * young = [E1, E2] (E(x) = (1 - x^p) E1 + x^p E2, problem/elasticity.hpp:13-17). */
typedef struct icarat_grid_desc {
  int32_t nx, ny, nz;     /* [mesh] division */
  double lx, ly, lz;      /* [mesh] length */
  double young1, young2;  /* [material] young */
  double poisson;         /* [material] poisson */
  double penalty;         /* [material] penalty */
} icarat_grid_desc;

typedef struct icarat_solve_opts {
  double tol;             /* relative residual ||f - K u|| / ||f|| to reach */
  int32_t maxit;          /* iteration cap of the (inner) Krylov method */
  int32_t precision;      /* ICARAT_HIP_FP64 / FP32 / BF16 */
  int32_t precond;        /* ICARAT_HIP_PRECOND_* */
  double inner_tol;       /* mixed precision: reliable-update ratio delta -- the FP64 true
                             residual replaces the FP32 recursive one whenever the latter
                             has dropped by delta since the last replacement (default 0.1) */
  int32_t max_refine;     /* mixed precision: cap on residual replacements (default 200) */
  int32_t warm;           /* 1: start the iteration from this load case's last solution on
                             this context (if any) instead of 0 -- same stop test; the
                             reference starts from 0 (FP64 Jacobi and multigrid PCG) */
} icarat_solve_opts;

typedef struct icarat_solve_stats {
  int32_t iters;          /* Krylov iterations (all refinement steps summed) */
  int32_t refine_steps;   /* FP64 residual replacements (reliable updates / verified stops) */
  double rec_relres;      /* recursive (Krylov) relative residual at exit */
  double true_relres;     /* FP64 ||f - K u|| / ||f|| recomputed by the FP64 kernel */
  double seconds;         /* device wall time of the solve */
  int32_t fallback;       /* learned preconditioner: 0 none, else it fell back to its base
                             method (Jacobi-PCG or multigrid PCG) from the current iterate
                             because the residual stalled (1), the iteration broke down (2)
                             or z.r <= 0 (3) */
  int32_t status;         /* ICARAT_HIP_OK or ICARAT_HIP_NOT_CONVERGED */
} icarat_solve_stats;

/* ---- context ---------------------------------------------------------- */
int icarat_hip_create(const icarat_grid_desc *desc, icarat_hip_ctx **out);
int icarat_hip_destroy(icarat_hip_ctx *ctx);
const char *icarat_hip_last_error(const icarat_hip_ctx *ctx);
/* numnp, nelx, neq = 3*numnp (io/mesh_icarat.hpp:106-111), numeq (-1 before make_dof) */
int icarat_hip_counts(const icarat_hip_ctx *ctx, int64_t *numnp, int64_t *nelx, int64_t *neq,
                      int32_t *numeq);
/* 1 if a HIP device is usable from this process, else 0 */
int icarat_hip_device_available(void);

/* ---- configuration and result files ------------------------------------- */
typedef struct icarat_toml_info {
  int32_t nx, ny, nz;      /* [mesh] division */
  double lx, ly, lz;       /* [mesh] length */
  double designs_0;        /* [material] designs_0 (1 if absent): the initial design */
  double filter_radius;    /* [this_problem] filterR (0 if absent) */
  int32_t n_dirichlet;     /* entries of [boundary_condition] dirichlet applied */
  int32_t n_neumann;       /* entries of [boundary_condition] neumann applied */
} icarat_toml_info;
/* A context from one of ICARAT's configuration files (toml::parse + MeshIcarat +
 * Elastic::setParameter + BCIcarat, io/mesh_icarat.hpp:72-139,
 * problem/elasticity.cpp:19-33, io/boundary_condition_icarat.hpp:72-264): [mesh]
 * dimension 3 / type "hexa8" / division / length, [material] young / poisson /
 * penalty / designs_0, [boundary_condition] dirichlet and neumann boxes (applied in
 * that order, with make_dof between them), [this_problem] filterR.  Other keys are
 * ignored; other element types are rejected.  info and err (err_len bytes, the
 * message of a failure) may be NULL. */
int icarat_hip_create_from_toml(const char *path, icarat_hip_ctx **out, icarat_toml_info *info,
                                char *err, int32_t err_len);
/* The drivers' result file (output(), opt_macro_oc.cpp:214-236, with the writer of
 * io/export_vtk.hpp:47-190): legacy ASCII VTK, nodes, hexa8 cells, and, when a
 * solution exists, POINT_DATA "displacement" and CELL_DATA "mises"; CELL_DATA
 * "density" when density (nelx) is not NULL. */
int icarat_hip_write_vtk(icarat_hip_ctx *ctx, const char *path, const double *density);

/* The drivers' history files (io/export_graph.hpp): addPlot(x, y, filename,
 * isFirst) appends the line "x y" (default stream formatting), creating the file
 * afresh when is_first == 1 (export_graph.hpp:56-70); exportGraph(xx, yy,
 * filename) writes n lines "xx[i] yy[i]", or "i yy[i]" when xx is NULL
 * (export_graph.hpp:15-40).  Host only. */
int icarat_hip_add_plot(const char *filename, double x, double y, int32_t is_first);
int icarat_hip_export_graph(const char *filename, const double *xx, const double *yy, int64_t n);

/* ---- boundary conditions (BCIcarat) -------------------------------------- */
/* Dirichlet box: flag[6]/val[6]/range {x0,x1,y0,y1,z0,z1}, tolerance 1e-7
 * (io/boundary_condition_icarat.hpp:72-138, io/boundary_condition_base.hpp:19-33). */
int icarat_hip_add_dirichlet(icarat_hip_ctx *ctx, const int32_t flag[6], const double val[6],
                             const double range[6]);
/* DOF numbering (BCIcarat::makeDOF -> countDOF, io/boundary_condition_icarat.hpp:140-148). */
int icarat_hip_make_dof(icarat_hip_ctx *ctx, int32_t *numeq);
/* Neumann box: point load if degenerate, else a face load integrated on every
 * element face with 4 nodes in the box (BCIcarat::input_neumann /
 * generateLoadMain, io/boundary_condition_icarat.hpp:212-410). */
int icarat_hip_add_neumann(icarat_hip_ctx *ctx, const double val[3], const double range[6]);
/* Load cases (one Force per case, reference sample/software/opt-multi-load/
 * opt_multi_load.cpp:74-84).  make_dof leaves one empty case (index 0);
 * new_load_case appends an empty case and makes it current.  add_neumann,
 * get_fext / set_fext and solve act on the current case; simp_run sums the
 * compliance and the sensitivities over all cases (opt_multi_load.cpp:118-146). */
int icarat_hip_new_load_case(icarat_hip_ctx *ctx, int32_t *index);
int icarat_hip_select_load_case(icarat_hip_ctx *ctx, int32_t index);
int icarat_hip_load_case_count(const icarat_hip_ctx *ctx, int32_t *n);
/* dof[3*node + c] = reference DOF number (Node::dof, problem/base.hpp:86) */
int icarat_hip_get_dof(const icarat_hip_ctx *ctx, int32_t *dof);
/* Force::fext_org (problem/base.hpp:131), numeq entries */
int icarat_hip_get_fext(const icarat_hip_ctx *ctx, double *f);
int icarat_hip_set_fext(icarat_hip_ctx *ctx, const double *f);
/* Shared element stiffness at young = 1 (24x24, row-major, DOF order 3a+c):
 * Ke of LinearElastic::makeKeFint (problem/elasticity.cpp:211-255) is young(x) * KE0. */
int icarat_hip_get_ke0(const icarat_hip_ctx *ctx, double *ke);

/* ---- analysis (LinearAnalysis / Solvers) --------------------------------- */
/* Physical densities (Elastic::setDesignS for every element), nelx entries. */
int icarat_hip_set_density(icarat_hip_ctx *ctx, const double *xphys, int64_t n);
/* y = K(x) u on free DOFs (the assembled globalK of LinearAnalysis::assemblingK,
 * analysis/linear.hpp:117-166, applied matrix-free).  precision selects the
 * contraction kernel (ICARAT_HIP_FP64 / FP32 / BF16). */
int icarat_hip_apply_K(icarat_hip_ctx *ctx, const double *u, double *y, int32_t precision);
/* diag(K) on free DOFs (DiagonalPreconditioner, analysis/solver/CG.cpp:59) */
int icarat_hip_get_diag(icarat_hip_ctx *ctx, double *d);
/* Equilibrium solve K u = fext with x0 = 0 (LinearAnalysis::solve +
 * Solvers::ScalingCG, analysis/linear.hpp:55-104, analysis/solver/CG.cpp:56-100).
 * u: numeq solution.  Returns ICARAT_HIP_NOT_CONVERGED when the cap is hit
 * (u then holds the last iterate). */
int icarat_hip_solve(icarat_hip_ctx *ctx, const icarat_solve_opts *opts, double *u,
                     icarat_solve_stats *stats);
/* Every load case of the context solved (the loop of
 * sample/software/opt-multi-load/opt_multi_load.cpp:118-146, which solves each
 * load case in turn).  With FP64 Jacobi-PCG or FP64 multigrid PCG (and the
 * z-marching K.u; for multigrid also no Galerkin coarse operators, an even
 * batch_iters, w_switch 0) all cases run as one multi-column iteration: each
 * launch applies K (or a cycle's transfer, sweep or dense solve) to every
 * case, each case keeps its own scalars, verified stop and endgame, and each
 * case's arithmetic is that of a solve of the case alone, so the results
 * equal icarat_hip_solve per case bit for bit.  Other preconditioners and
 * precisions solve the cases one after another.  u: ncase x numeq (may be
 * NULL); stats: ncase entries (may be NULL).  Afterwards the context holds the
 * selected case's solution (compliance, sensitivity, mises). */
int icarat_hip_solve_cases(icarat_hip_ctx *ctx, const icarat_solve_opts *opts, double *u,
                           icarat_solve_stats *stats);
/* ||fext - K u|| / ||fext|| in FP64 (analysis/linear.hpp:93-103). */
int icarat_hip_true_residual(icarat_hip_ctx *ctx, const double *u, double *relres);
/* Compliance c = u . fext of the last solve (opt_macro_oc.cpp:110) and the
 * element sensitivities of the driver's sensitivity() (opt_macro_oc.cpp:182-210):
 * dfds = -(E'/E) u_e^T Ke u_e, dgds = volume / vnorm.  Either output may be NULL. */
int icarat_hip_compliance_sensitivity(icarat_hip_ctx *ctx, double vnorm, double *c, double *dfds,
                                      double *dgds);

/* Element-mean stress and von Mises stress of the last solve
 * (Elastic::makeMisesStress, problem/elasticity.cpp:166-209): stress = mean over
 * the Gauss points of De(E(x_e)) B u_e, mises = sqrt(stress^T V stress).
 * mises: nelx entries; stress (may be NULL): nelx x 6, Voigt [xx,yy,zz,xy,yz,zx]. */
int icarat_hip_mises(icarat_hip_ctx *ctx, double *mises, double *stress);
/* No reference counterpart: icarat_hip_mises for callers that cannot allocate
 * device memory (code under graph capture, tight serving loops).  The same
 * outputs; the element stresses are staged in a fixed device array of the
 * library (48 MiB per GPU, held until the process ends) instead of a per-call
 * allocation, so the call allocates nothing.  With stress != NULL the context
 * may have at most 1,048,576 elements (else ICARAT_HIP_INVALID_ARGUMENT).  All
 * calls on one GPU share that array: calls that overlap on one GPU race. */
int icarat_hip_mises_noalloc(icarat_hip_ctx *ctx, double *mises, double *stress);

/* ---- optimization (Filter / OC / MMASolver) ------------------------------ */
/* Filter<E,N>(fem, element, node, radius) (optimization/filter.hpp:63-70). */
int icarat_hip_filter_setup(icarat_hip_ctx *ctx, double radius);
/* Filter::densityFilter(in, is_sens) (optimization/filter.hpp:106-155). */
int icarat_hip_filter(icarat_hip_ctx *ctx, const double *in, double *out, int32_t is_sens);
/* Filter::sensitivityFilter(design_s, dfds) (optimization/filter.hpp:76-104):
 * Sigmund's sensitivity filter over the cone neighbours of icarat_hip_filter_setup,
 * out_i = (x_i V)/(x_i sum_j w_ij) * sum_j (w_ij x_j dfds_j)/(x_j V); x, dfds, out
 * hold nelx values (host memory).  x must be nonzero (the reference divides by it). */
int icarat_hip_sensitivity_filter(icarat_hip_ctx *ctx, const double *x, const double *dfds,
                                  double *out);
/* Filter::threshold (deriv = 0) / thresholdDerivative (deriv = 1)
 * (optimization/filter.hpp:216-248). */
int icarat_hip_threshold(icarat_hip_ctx *ctx, const double *in, double beta, double T,
                         double *out, int32_t deriv);
/* FilterHelmholtz<E,N>(fem, element, node, radius) (optimization/filter_helm.hpp:69-75):
 * the Helmholtz PDE filter on the voxel box, natural boundary conditions. */
int icarat_hip_helm_setup(icarat_hip_ctx *ctx, double radius);
/* FilterHelmholtz::densityFilter(design_s, solvName, log, sensFlag)
 * (optimization/filter_helm.hpp:76-135, 191-257): is_sens = sensFlag.  The
 * reference assembles the nodal system sum_e (N N^T + (r/(2 sqrt 3))^2 B^T B) jac
 * and solves it with the named iterative solver ("Eigen_CG", relative residual
 * 1e-10); here the same system is solved directly by fast diagonalisation (the
 * separable voxel operator, FP64 GEMMs), so the result is the exact solution to
 * FP64 rounding.  threshold / thresholdDerivative are icarat_hip_threshold. */
int icarat_hip_helm_filter(icarat_hip_ctx *ctx, const double *in, double *out, int32_t is_sens);
/* OC::UpdateVariables with the drivers' volume callback (filter, optional
 * projection, sum * volume) (optimization/opt-solver/OCSolver.hpp:20-79,
 * opt_macro_oc.cpp:153-164).  x updated in place; *steps = bisection steps. */
int icarat_hip_oc_update(icarat_hip_ctx *ctx, double *x, const double *dfdx, double g,
                         const double *dgdx, double xmin, double xmax, int32_t use_projection,
                         double beta, int32_t *steps);

/* MMASolver(n = nelx, m = 1) (optimization/opt-solver/MMASolver.hpp:62-116) on the
 * device: SetMoveLimit / SetOscillation / SetAsymptotes
 * (MMASolver.hpp:72-75) and Reset. */
int icarat_hip_mma_set_params(icarat_hip_ctx *ctx, double move, double albefa, double asyminit,
                              double asymdec, double asyminc);
int icarat_hip_mma_reset(icarat_hip_ctx *ctx);
/* MMASolver::Update(x, dfdx, &g, dgdx, xmin, xmax) for the single volume
 * constraint (optimization/opt-solver/MMASolver.cpp:66-82); x updated in place;
 * *outer = interior-point outer loops (may be NULL). */
int icarat_hip_mma_update(icarat_hip_ctx *ctx, double *x, const double *dfdx, double g,
                          const double *dgdx, double xmin, double xmax, int32_t *outer);

/* MMASolver(n, m, a, c, d) (optimization/opt-solver/MMASolver.hpp:62-116,
 * MMASolver.cpp) for any n and 1 <= m <= 8 constraints, on the device with its
 * own stream; not tied to a context.  The subproblem (GenSub), x(lambda) and the
 * dual sums run as element passes, the dual interior-point iteration's m x m
 * Newton system (Factorize / Solve: LU without pivoting, as the reference), line
 * search and epsilon loop in one device block.  set_params = SetMoveLimit /
 * SetOscillation / SetAsymptotes (MMASolver.hpp:68-74); reset = Reset; update =
 * Update(xval, dfdx, gx, dgdx, xmin, xmax) (MMASolver.cpp:66-82) with host arrays:
 * x (n, updated in place), dfdx (n), gx (m), dgdx (n m, dgdx[i m + j]), xmin and
 * xmax (n); *outer = interior-point outer loops (may be NULL).  asymptotes:
 * the last update's low and upp (n each, either may be NULL). */
typedef struct icarat_hip_mmasolver icarat_hip_mmasolver;
int icarat_hip_mmasolver_create(int32_t n, int32_t m, double a, double c, double d,
                                icarat_hip_mmasolver **out);
int icarat_hip_mmasolver_destroy(icarat_hip_mmasolver *h);
const char *icarat_hip_mmasolver_last_error(const icarat_hip_mmasolver *h);
int icarat_hip_mmasolver_set_params(icarat_hip_mmasolver *h, double move, double albefa,
                                    double asyminit, double asymdec, double asyminc);
int icarat_hip_mmasolver_reset(icarat_hip_mmasolver *h);
int icarat_hip_mmasolver_update(icarat_hip_mmasolver *h, double *x, const double *dfdx,
                                const double *gx, const double *dgdx, const double *xmin,
                                const double *xmax, int32_t *outer);
int icarat_hip_mmasolver_asymptotes(icarat_hip_mmasolver *h, double *low, double *upp);

/* GCMMASolver(n, m, a, c, d) (optimization/opt-solver/GCMMASolver.hpp:62-85,
 * GCMMASolver.cpp) for any n and 1 <= m <= 8 constraints, on the device with its
 * own stream; not tied to a context.  The subproblem, the approximation sums,
 * raacof and the dual iteration (as icarat_hip_mmasolver) run on the device; raa0,
 * raa, the approximation constants and ConCheck on the host.  set_params =
 * SetMoveLimit / SetOscillation / SetAsymptotes; reset = Reset (the outer counter);
 * outer = OuterUpdate(xmma, xval, f0x, df0dx, fx, dfdx, xmin, xmax); inner =
 * InnerUpdate(xmma, f0xnew, fxnew, xval, f0x, df0dx, fx, dfdx, xmin, xmax) with xmma
 * in/out; concheck = ConCheck(f0xnew, fxnew) into *ok; host arrays: xmma, xval,
 * df0dx, xmin, xmax (n), fx, fxnew (m), dfdx (n m, dfdx[i m + j]).  state: {raa0,
 * raa[m], f0app, fapp[m]} (2 + 2 m values). */
typedef struct icarat_hip_gcmmasolver icarat_hip_gcmmasolver;
int icarat_hip_gcmmasolver_create(int32_t n, int32_t m, double a, double c, double d,
                                  icarat_hip_gcmmasolver **out);
int icarat_hip_gcmmasolver_destroy(icarat_hip_gcmmasolver *h);
const char *icarat_hip_gcmmasolver_last_error(const icarat_hip_gcmmasolver *h);
int icarat_hip_gcmmasolver_set_params(icarat_hip_gcmmasolver *h, double move, double albefa,
                                      double asyminit, double asymdec, double asyminc);
int icarat_hip_gcmmasolver_reset(icarat_hip_gcmmasolver *h);
int icarat_hip_gcmmasolver_outer(icarat_hip_gcmmasolver *h, double *xmma, const double *xval,
                                 double f0x, const double *df0dx, const double *fx,
                                 const double *dfdx, const double *xmin, const double *xmax);
int icarat_hip_gcmmasolver_inner(icarat_hip_gcmmasolver *h, double *xmma, double f0xnew,
                                 const double *fxnew, const double *xval, double f0x,
                                 const double *df0dx, const double *fx, const double *dfdx,
                                 const double *xmin, const double *xmax);
int icarat_hip_gcmmasolver_concheck(icarat_hip_gcmmasolver *h, double f0xnew, const double *fxnew,
                                    int32_t *ok);
int icarat_hip_gcmmasolver_state(icarat_hip_gcmmasolver *h, double *out);

/* GCMMASolver(n = nelx, m = 1, a = 0, c = 1000, d = 0) (optimization/opt-solver/
 * GCMMASolver.hpp:62-85) on the device, for the single volume constraint fx:
 * SetMoveLimit / SetOscillation / SetAsymptotes, and Reset (which here also
 * restores raa0 = raa = 1).  outer = OuterUpdate(xmma, xval, f0x, df0dx, &fx,
 * dfdx, xmin, xmax); inner = InnerUpdate(xmma, f0xnew, &fxnew, xval, f0x, df0dx,
 * &fx, dfdx, xmin, xmax) with xmma in/out; concheck = ConCheck(f0xnew, &fxnew).
 * Bounds are uniform. */
int icarat_hip_gcmma_set_params(icarat_hip_ctx *ctx, double move, double albefa,
                                double asyminit, double asymdec, double asyminc);
int icarat_hip_gcmma_reset(icarat_hip_ctx *ctx);
int icarat_hip_gcmma_outer(icarat_hip_ctx *ctx, double *xmma, const double *xval, double f0x,
                           const double *df0dx, double fx, const double *dfdx, double xmin,
                           double xmax);
int icarat_hip_gcmma_inner(icarat_hip_ctx *ctx, double *xmma, double f0xnew, double fxnew,
                           const double *xval, double f0x, const double *df0dx, double fx,
                           const double *dfdx, double xmin, double xmax);
int icarat_hip_gcmma_concheck(icarat_hip_ctx *ctx, double f0xnew, double fxnew,
                              int32_t *conservative);

/* ---- SIMP driver loop ------------------------------------------------------ */
typedef struct icarat_simp_opts {
  int32_t n_iter;          /* design iterations */
  int32_t optimizer;       /* 0 = OC, 1 = MMA (m = 1 volume constraint), 2 = GCMMA (the loop of
                              opt-macro-gcmma/opt_macro_gcmma.cpp:84-212: projection beta =
                              min(int(max((k - 100) / 40, 1)), 8) when use_projection, constraint
                              volume - V0, inner iterations until conservative; icarat_hip_simp_run
                              only) */
  int32_t use_projection;  /* 0 = plain method (no tanh projection) */
  int32_t filter;          /* 0 = cone filter (icarat_hip_filter_setup, optimization/filter.hpp),
                              1 = Helmholtz filter (icarat_hip_helm_setup, filter_helm.hpp; the
                              loop of opt-mma-HelmholtzFilter/opt_mma_Helmholtz.cpp:84-150) */
  double xmin, xmax;       /* Optimize(nelx, 1, xmin, xmax) (optimization/opt_base.hpp:20-24) */
  icarat_solve_opts solve;
  double conv_eps;         /* Optimize::isConvergence(conv_eps, conv_minstep) (opt_base.hpp:28-44,
                              called after the compliance, opt_macro_oc.cpp:137): the loop ends
                              with design iteration k (its update included) once
                              |c_old - c_k| / c_k < conv_eps and k > conv_minstep, c_old being the
                              compliance of the last iteration that did not pass; n_iter stays the
                              cap.  0: always n_iter iterations (OC and MMA loops) */
  int32_t conv_minstep;
  const char *history_csv; /* addPlot(step, compliance, history_csv, step) every iteration
                              (io/export_graph.hpp:56-70, opt_macro_oc.cpp:134); NULL: none */
  int32_t case_mode;       /* several load cases: 0 = automatic (solved together as in
                              icarat_hip_solve_cases where the solve options allow it: multigrid
                              PCG at every size, Jacobi-PCG up to 600,000 elements, above which
                              one launch per case was faster), 1 = one after another, 2 =
                              together where the options allow it.  The results are the same in
                              every mode; icarat_hip_batch_simp_run follows the same choice for
                              multigrid PCG and solves Jacobi-PCG cases one after another */
} icarat_simp_opts;

typedef struct icarat_simp_stats {
  double seconds;          /* wall time of the whole loop */
  double solve_s, filter_s, sens_s, update_s;
  int64_t pcg_iters;       /* Krylov iterations, all design iterations */
  double max_true_relres;  /* largest FP64 true residual over the loop */
  double train_s;          /* learned preconditioner: time spent training the network */
  int64_t fallbacks;       /* learned preconditioner: solves that fell back to Jacobi-PCG */
  int64_t inner_iters;     /* GCMMA: inner iterations, all design iterations */
  int32_t steps;           /* design iterations run */
  int32_t converged;       /* 1 if the loop ended on the conv_eps test */
} icarat_simp_stats;

/* The reference drivers' loop (sample/algorithm/opt-macro-oc/opt_macro_oc.cpp:91-178,
 * opt-macro-mma-nondesign/opt_macro_mma_nondesign.cpp:90-178) with the cone
 * filter set up by icarat_hip_filter_setup, or the Helmholtz filter set up by
 * icarat_hip_helm_setup (opts->filter = 1).  x: design variables in/out
 * (nelx).  hist (may be NULL): per iteration {compliance, volume, iterations,
 * true relative residual}. */
int icarat_hip_simp_run(icarat_hip_ctx *ctx, const icarat_simp_opts *opts, double *x,
                        double *hist, icarat_simp_stats *stats);

/* ---- learned preconditioner --------------------------------------------------
 * No reference counterpart (the reference solves with Scaling_CG only,
 * analysis/solver/CG.cpp:56-100): a 3-level U-Net on a cell grid coarsened 2x
 * from the node lattice corrects the Jacobi step inside flexible PCG,
 *     z = D^-1 r + s P(net(log-stiffness, R(D^-1 r) / s)),  s = RMS of R(D^-1 r),
 * R full-weighting restriction, P its transpose.  The network is trained on
 * (residual, error) pairs harvested from Jacobi-PCG solves.  A solve with
 * precond = ICARAT_HIP_PRECOND_LEARNED falls back to Jacobi-PCG from its
 * current iterate when z.r <= 0 or the residual stalls. */
typedef struct icarat_learned_opts {
  int32_t train_iters;     /* simp_run: the first train_iters solves of the preconditioner's
                              lifetime run Jacobi-PCG and are harvested */
  int32_t train_steps;     /* simp_run: Adam steps in total, spread over the harvested iterations */
  int32_t snapshot_every;  /* PCG iterations between two training snapshots of one solve */
  int32_t max_samples;     /* capacity of the sample ring */
  int32_t batch_iters;     /* flexible-PCG iterations per graph launch */
  int32_t stall_iters;     /* fallback when the residual has not halved for this many iterations */
  float lr;                /* Adam learning rate */
  float gate;              /* the network is used only while its mean relative squared error
                              on the stored samples is below gate (else z = D^-1 r) */
  uint64_t seed;           /* weight initialisation and sample order */
  int32_t base;            /* the preconditioner the network corrects: 0 Jacobi,
                              z = D^-1 r + s P(net(feature, R(D^-1 r) / s)); 1 the multigrid
                              V-cycle of icarat_hip_mg_setup, z = V(r) + s P(net(feature,
                              R(V(r)) / s)), trained on the error the V-cycle leaves, e - V(r),
                              from multigrid-PCG iterates; the fallback is then multigrid PCG */
  float correction_scale;  /* the network's correction is added times this (<= 0: 1) */
} icarat_learned_opts;

typedef struct icarat_learned_info {
  int32_t samples;         /* (residual, error) pairs held */
  int32_t active;          /* 1 if the network passed the gate and is used in solves */
  int64_t steps;           /* Adam steps taken */
  double loss;             /* mean relative squared error of the last training steps */
  double val_loss;         /* mean relative squared error over all stored samples */
  double gain;             /* output scale: median RMS of the targets in residual units */
  int64_t n_params;        /* network parameters */
  int32_t grid[3];         /* cell grid X, Y, Z */
  int32_t harvested;       /* harvesting solves finished */
  int32_t fallback_iter;   /* iteration at which the last solve fell back (-1: none yet) */
  int32_t trial_iters;     /* simp_run: iterations of the first solve with the network (-1: none) */
  int32_t trial_base_iters;/* iterations of the last harvesting solve it was held against; a
                              trial that needs more switches the network off for the loop */
} icarat_learned_info;

/* Fills *o with the defaults (10, 400, 64, 160, 8, 400, 1e-3, 0.8, 1, 0, 1). */
int icarat_hip_learned_default_opts(icarat_learned_opts *o);
/* (Re)creates the preconditioner of the context (o = NULL: defaults).  Needs make_dof. */
int icarat_hip_learned_setup(icarat_hip_ctx *ctx, const icarat_learned_opts *o);
/* The base method's solve of the current density (as icarat_hip_solve with
 * precond = JACOBI, or MG for the multigrid base) that stores training
 * snapshots; u may be NULL. */
int icarat_hip_learned_harvest(icarat_hip_ctx *ctx, const icarat_solve_opts *opts, double *u,
                               icarat_solve_stats *stats);
/* Adam steps on the stored samples, then the gate on all samples; *loss = mean
 * ||net - t||^2 / ||t||^2 of the last steps (-1 without samples). */
int icarat_hip_learned_train(icarat_hip_ctx *ctx, int32_t steps, double *loss);
int icarat_hip_learned_info(const icarat_hip_ctx *ctx, icarat_learned_info *info);
/* Network parameters in the layout of icarat_hip_unet_get_params.  Setting
 * them switches the network on (gain 1 until a harvest sets the gain). */
int icarat_hip_learned_get_params(const icarat_hip_ctx *ctx, float *params);
int icarat_hip_learned_set_params(icarat_hip_ctx *ctx, const float *params);
/* z = M(r) for a numeq residual r (density as last set). */
int icarat_hip_learned_apply(icarat_hip_ctx *ctx, const double *r, double *z);

/* The network alone on a cell grid (X, Y, Z multiples of 4), for checks of
 * its kernels.  Activations are channels-last with x fastest: x[v*4 + c]
 * (4 input channels; the kernels pad to 8 with zeros), out[v*3 + c].
 * Parameters: per layer c0..c5 the weights W[tap][co][ci] (tap = 9(dz+1) +
 * 3(dy+1) + (dx+1); 27 taps, 1 for c5) then the biases b[co]; channels
 * (cin, cout) = (8,16) (16,32) (32,64) (96,32) (48,16) (16,3). */
typedef struct icarat_hip_unet icarat_hip_unet;
int icarat_hip_unet_create(int32_t X, int32_t Y, int32_t Z, uint64_t seed, icarat_hip_unet **out);
int icarat_hip_unet_destroy(icarat_hip_unet *net);
const char *icarat_hip_unet_last_error(const icarat_hip_unet *net);
int icarat_hip_unet_param_count(int64_t *n);
int icarat_hip_unet_set_params(icarat_hip_unet *net, const float *params);
int icarat_hip_unet_get_params(const icarat_hip_unet *net, float *params);
int icarat_hip_unet_forward(icarat_hip_unet *net, const float *x, float *out);
/* loss = 0.5 ||out - t||^2 / ||t||^2 and its parameter gradient (no update) */
int icarat_hip_unet_grads(icarat_hip_unet *net, const float *x, const float *target,
                          double *loss, float *grads);
/* one Adam step (b1 0.9, b2 0.999, eps 1e-8) on the loss above; *loss before the step */
int icarat_hip_unet_train_step(icarat_hip_unet *net, const float *x, const float *target,
                               float lr, double *loss);

/* ---- batched problems ------------------------------------------------------------
 * No reference counterpart (the reference runs one problem per process): several
 * independent problems of any size, each a prepared context (mesh, supports,
 * load cases, filter), solved together.  Per launch every running problem's own
 * captured graph (a batch of its iterations) goes on its context's stream, forked
 * from and joined into the batch stream, so the problems run concurrently on
 * separate hardware queues; each problem keeps its own
 * FP64 verified stop.  Batched solves are FP64 Jacobi-PCG or FP64 multigrid PCG
 * (opts precond ICARAT_HIP_PRECOND_MG; each problem's hierarchy is set up with
 * the defaults unless icarat_hip_mg_setup was called on its context).  A context
 * added to a batch must outlive it and must not be used on its own while the
 * batch runs. */
typedef struct icarat_hip_batch icarat_hip_batch;
int icarat_hip_batch_create(icarat_hip_batch **out);
int icarat_hip_batch_destroy(icarat_hip_batch *b);
const char *icarat_hip_batch_last_error(const icarat_hip_batch *b);
int icarat_hip_batch_add(icarat_hip_batch *b, icarat_hip_ctx *ctx);
/* Admission control: problems join the running set largest first while the
 * running problems' elements stay within `elements` (at least one runs; 0: all
 * run at once); an admitted problem runs its whole loop.  Default 0. */
int icarat_hip_batch_set_budget(icarat_hip_batch *b, int64_t elements);
/* problems in the batch and launches (one graph per running problem each) so far */
int icarat_hip_batch_info(const icarat_hip_batch *b, int32_t *n, int64_t *graph_launches);
/* one solve of every problem's current load case; stats[n] (may be NULL) */
int icarat_hip_batch_solve(icarat_hip_batch *b, const icarat_solve_opts *opts,
                           icarat_solve_stats *stats);
/* the SIMP loop of every problem in lock step (as icarat_hip_simp_run each, load
 * cases summed): x[i] (nelx of problem i) in/out, hist[i] (may be NULL, or hist
 * itself NULL) 4 * n_iter per problem, stats[n] (may be NULL) */
int icarat_hip_batch_simp_run(icarat_hip_batch *b, const icarat_simp_opts *opts,
                              double *const *x, double *const *hist, icarat_simp_stats *stats);
/* Autotuner: times reps whole-batch launches (a batch of Jacobi-PCG iterations of
 * every problem) under each stencil configuration -- 0..2: z-marching columns with
 * the default z-chunk count divided by 1, 2, 4; 3: 3-D halo tiles -- and keeps
 * the fastest for every problem.  Needs densities on every problem; overwrites
 * the solver state.  *config = the chosen configuration, ms[4] (may be NULL) the
 * time of one launch per configuration. */
int icarat_hip_batch_autotune(icarat_hip_batch *b, int32_t reps, int32_t *config, double *ms);

/* ---- multigrid preconditioner --------------------------------------------------
 * No reference counterpart (the reference solves with Jacobi-PCG): a hierarchy of
 * 2x coarser voxel grids of the same box (coarse Young's moduli = mean of the 8
 * children; the Dirichlet boxes replayed, so coarse constraints sit at the same
 * coordinates; coarsening stops at an odd dimension), damped Jacobi smoothing
 * with each level's element stiffness, trilinear transfers, and a symmetric
 * V-cycle as the preconditioner of PCG.  The FP64 verified stop is unchanged.
 * A solve with precond = ICARAT_HIP_PRECOND_MG sets the hierarchy up with the
 * defaults when icarat_hip_mg_setup was not called. */
typedef struct icarat_mg_opts {
  int32_t max_levels;      /* levels including the fine grid */
  int32_t nu_pre, nu_post; /* damped Jacobi sweeps before (>= 0) / after (>= 1) the coarse
                              correction; nu_pre != nu_post gives a non-symmetric V-cycle,
                              which the flexible PCG accepts */
  int32_t nu_coarse;       /* sweeps on the coarsest grid (unless it is solved exactly) */
  double omega;            /* Jacobi damping relative to 1 / mu, mu = lambda_max(diag(KE0)^-1
                              KE0) >= lambda_max(D^-1 K) for any densities; in (0, 2) */
  int32_t batch_iters;     /* PCG iterations per graph launch */
  int32_t dense_dofs;      /* coarsening stops at the first grid of at most this many lattice
                              DOFs (3 x nodes, capped at 6144), which is solved exactly by
                              its dense inverse; a coarsest grid that cannot be halved
                              further is solved exactly up to 1.5 x this many (0: never;
                              the coarsest grid is smoothed; -1: automatic, 1/3000 of the
                              fine lattice DOFs, at least 1024) */
  int32_t direct_nodes;    /* levels of at most this many nodes smooth with the fused
                              one-thread-per-node sweep; larger ones with the z-marching K.u */
  int32_t coarse_avg;      /* coarse operators: a coarse element's Young's modulus from its 8
                              children (0 arithmetic mean, 1 geometric mean, 2 the SIMP law of
                              the mean density, 3 quadratic mean, 4 maximum) times the coarse
                              KE0; or 5: Galerkin, R A P element by element (explicit 24x24
                              matrices per coarse element, 4.6 KB each) */
  int32_t w_levels;        /* coarse levels 1..w_levels are cycled twice per visit (a W-cycle
                              at the top of the hierarchy; 0: V-cycle; < 0: automatic by the
                              fine mesh: 0 below 500,000 elements, 1 up to 2,000,000, every
                              coarse level above) */
  double w_switch;         /* a solve's first phase -- the W-cycle (w_levels > 0) and the
                              FP32 fine level (precision FP32) -- ends once its recursive
                              relative residual is below w_switch x tol (0: the W-cycle runs
                              to the end); it goes on with the FP64 V-cycle (restarted with
                              beta = 0).  The FP32 fine level ends at 30 x tol at the latest:
                              its sweeps cannot resolve the last digits of a field at its FP64
                              floor */
  int32_t compensated;     /* 1: the PCG iterate accumulates its updates as a compensated sum
                              (x plus a low-part vector, x its rounding), so updates below
                              half an ulp of x -- the large displacements of near-void
                              regions near the tolerance -- are not lost while the recursive
                              residual counts them; 0: x += alpha p */
} icarat_mg_opts;
/* Fills *o with the defaults (10, 1, 1, 6, 1.95, 4, -1, 262144, 0, -1, 0, 1). */
int icarat_hip_mg_default_opts(icarat_mg_opts *o);
/* (Re)builds the hierarchy of the context (o = NULL: defaults).  Needs make_dof. */
int icarat_hip_mg_setup(icarat_hip_ctx *ctx, const icarat_mg_opts *o);
/* levels of the hierarchy (1 = no coarsening possible); dims (may be NULL) gets
 * nx, ny, nz of every level, 3 * levels entries */
int icarat_hip_mg_info(const icarat_hip_ctx *ctx, int32_t *levels, int32_t *dims);
typedef struct icarat_mg_status {
  int32_t levels;          /* levels of the hierarchy */
  int32_t dense;           /* 1 if the coarsest level is solved exactly by its dense inverse */
  int32_t direct_fine;     /* 1 if the fine level smooths with the one-thread-per-node kernel */
  int32_t fine_fp32;       /* 1 if the last solve ran the FP32 fine level (precision FP32; in
                              its first phase) */
  int64_t factorizations;  /* Cholesky factorizations of the coarsest matrix so far */
  int64_t shifted;         /* of which broke down and were redone with the diagonal of the
                              free DOFs raised by a relative 1e-10 */
  int32_t w_levels;        /* coarse levels cycled twice (W-cycle phase of a solve; 0: V) */
  int64_t w_fallbacks;     /* W-cycle phases that broke down (z.r <= 0) and went on with the
                              V-cycle */
} icarat_mg_status;
/* State of the context's hierarchy (NOT_READY without one). */
int icarat_hip_mg_status(const icarat_hip_ctx *ctx, icarat_mg_status *status);

/* ---- measurement ------------------------------------------------------------ */
/* No reference counterpart: a timing hook for benchmarks.  Launches one kernel
 * `reps` times back to back on the context's stream between two HIP events
 * and returns the average duration and the algorithmic work of one launch
 * (flops and compulsory HBM bytes as documented for each kernel):
 *   kernel 0: FP64 PCG stencil half as the solver runs it (beta step reduced from the
 *             previous update's partials, p update + K.p + p.q), 1152 flop/element
 *   kernel 1: FP64 PCG vector half as the solver runs it (alpha step reduced from the
 *             stencil's partials, x, r update, r.z, r.r)
 *   kernel 2: density filter (isSens = false) at the set-up radius
 *   kernel 3: as 0 with KE0 read from memory (no symmetry classes)
 *   kernel 4: as 0 with the other FP64 kernel shape (see icarat_hip_set_stencil)
 *   kernel 5: PCG stencil half with the FP32 MFMA contraction (FP64 vectors)
 *   kernel 6: PCG stencil half with the BF16 MFMA contraction (FP64 vectors)
 *   kernel 7: plain copy of one lattice vector (3 Np doubles; bandwidth reference)
 *   kernel 8: FP64 y = K x as the multigrid smoother runs it (no PCG prologue),
 *             1152 flop/element
 *   kernel 9: one Helmholtz filter pass (icarat_hip_helm_setup): 4 numnp (NX+NY+NZ)
 *             flop of 1-D transforms (FP64 GEMMs), 16 B per element in and out
 *   kernel 10: FP64 PCG stencil of the multigrid solve (p = z + beta p_old from the
 *             V-cycle's z, no D^-1; q = K p, p.q), 1152 flop/element
 *   kernel 11: FP64 vector update of the multigrid solve (alpha step reduced from the
 *             stencil's partials, x, r update, r.r, z0 = omega D^-1 r)
 *   kernel 12: FP64 y = K x on the matrix cores (v_mfma_f64_4x4x4 element contraction,
 *             ICARAT_HIP_FP64_MFMA of icarat_hip_apply_K), 1152 flop/element
 * Needs set_density (and filter_setup for kernel 2, helm_setup for kernel 9).  Overwrites the solver's
 * work vectors. */
int icarat_hip_time_kernel(icarat_hip_ctx *ctx, int32_t kernel, int32_t reps, double *avg_ms,
                           double *flops, double *bytes);

/* No reference counterpart: selects the FP64 K.u kernel of this context.
 * kernel: 0 = z-marching node columns (default), 1 = 3-D halo tiles.
 * sym_max: highest KE0 symmetry class the kernel may use: 0 = KE0 read from
 * memory, 1 = box orbits (48 values in registers), 2 = cube orbits (12 values,
 * default); the class actually used is the highest one that reproduces KE0.
 * *sym_used (may be NULL) returns it. */
int icarat_hip_set_stencil(icarat_hip_ctx *ctx, int32_t kernel, int32_t sym_max, int32_t *sym_used);

/* ---- inputs ---------------------------------------------------------------- */
/* Seeded smooth density field (sum of nblob isotropic Gaussians, amplitude
 * U[0.5,1.5], centre U(domain), sigma U[smin,smax], rescaled to `mean`,
 * clipped to [lo,hi]).  Host only; returns the mean after clipping in *mean_out.
 * No reference counterpart: the input generator of the benchmark configs. */
int icarat_hip_gaussian_field(int32_t nx, int32_t ny, int32_t nz, double h, uint64_t seed,
                              int32_t nblob, double smin, double smax, double mean, double lo,
                              double hi, double *out, double *mean_out);

#ifdef __cplusplus
}
#endif
#endif /* ICARAT_HIP_H */
