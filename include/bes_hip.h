/* bes_hip.h — C ABI of the MI355X-native BES solver core (libbes_hip.so).
 *
 * The reference (joel-ho/BES) has no FFI: its residual and implicit-update
 * path is a set of abstract C++ classes called in-process.  Each entry point
 * below replaces one of those interfaces (cited as reference file:line); the
 * C++ host classes in include/bes/ and the Python bindings in bes_amd/ sit on
 * top of this ABI.  Plain pointers and sizes only; every function returns a
 * bes_status; nothing ever exits the process (the reference calls exit() on
 * errors, e.g. SolverExplicit.cpp:311-323).
 *
 * Conventions (reference Config.cpp:469-477, Face.hpp:33):
 *   conserved state  [rho, rho*u_d..., rho*E], nv = ndim + 2
 *   primitive state  [rho, u_d..., p, T],      nv + 1 entries
 *   host state arrays are cell-major [ncell][nv] in the reference's element
 *   order (the order of the mesh file); the device keeps its own SoA layout.
 *   face normals are unit, owner -> neighbour, outward on boundaries.
 */
#ifndef BES_HIP_H
#define BES_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  BES_OK = 0,
  BES_ERR_ARG = 1,     /* invalid argument / size / mode */
  BES_ERR_HIP = 2,     /* HIP runtime error (message in bes_last_error_message) */
  BES_ERR_NAN = 3,     /* non-finite flux; face index via bes_last_error_face */
  BES_ERR_MESH = 4,    /* mesh file unreadable or outside the supported subset */
  BES_ERR_SOLVE = 5    /* linear solver failed to reach its tolerance */
} bes_status;

/* boundary type codes, reference Options.hpp:34-35 */
enum { BES_INTERNAL = 1, BES_SLIP_WALL = 2, BES_FREESTREAM = 3 };

/* geometry mode bits (0 = the reference's metrics exactly):
 *   1: true tetrahedron volume  (reference MeshGmsh.cpp:101-133 omits the 1/3)
 *   2: true tetrahedron centroid (reference MeshGmsh.cpp:135-145 counts v2 twice)
 *   4: per-face reconstruction vector (reference Element.cpp:102-105 keys its face
 *      map before face ids exist, so Solver.cpp:91 reads the vector of local face 0) */
enum { BES_GEOM_COMPAT = 0, BES_GEOM_EXACT = 7 };

typedef struct {
  double gamma;      /* SPECIFIC_HEAT_RATIO, default 1.4     (Config.cpp:110) */
  double R;          /* GAS_CONSTANT, default 287             (Config.cpp:109) */
  double harten;     /* HARTEN_CORRECTION_LIMIT, default 0.05 (Config.cpp:107) */
  double venkat_k;   /* VENKATAKRISHNAN_FACTOR, default 5     (Config.cpp:106) */
} bes_gas;

/* Mesh in memory: the element lists of an MSH2 file (Mesh.cpp:95-558).
 * bfaces: boundary faces (lines in 2D, triangles in 3D) in file order, before
 * all cells; cells: triangles (CCW) or tetrahedra (positive), file order. */
typedef struct {
  int ndim;
  long nnode;
  const double *coords;   /* [nnode][ndim] */
  long ncell;
  const long *cells;      /* [ncell][ndim+1], 0-based */
  long nbface;
  const long *bfaces;     /* [nbface][ndim], 0-based */
  const int *btype;       /* [nbface] BES_SLIP_WALL / BES_FREESTREAM */
} bes_mesh_desc;

typedef struct bes_ctx bes_ctx;

/* ---- context: replaces Mesh::initialize + Solution/SolutionSnapshot storage
 *      (Mesh.cpp:66-93, Solution.cpp:43-61, SolutionSnapshot.cpp:60-86) ---- */
int bes_ctx_create(const bes_mesh_desc *mesh, int geom_mode, const bes_gas *gas, bes_ctx **out);
/* reads the reference's MSH2 subset (Mesh.cpp:95-558) */
int bes_ctx_create_msh2(const char *path, int geom_mode, const bes_gas *gas, bes_ctx **out);
int bes_ctx_destroy(bes_ctx *ctx);
int bes_ctx_info(const bes_ctx *ctx, int *ndim, long *ncell, long *nface);
/* reference-order geometry as the reference computes it (for parity and the
 * per-element getters); any pointer may be NULL.  Face arrays [nface][...],
 * cell arrays [ncell][...], k = ndim+1 faces per cell. */
int bes_get_geometry(const bes_ctx *ctx, long *owner, long *nbr, int *btype, double *S,
                     double *n, double *area, double *vol, double *centroid,
                     long *cell_faces, long *cell_adj, double *w, double *r);

/* freestream primitive [nv+1] used by far-field faces (Config.cpp:479-488,688-706);
 * sets every condition of the batch */
int bes_set_freestream(bes_ctx *ctx, const double *primitive);

/* ---- batched conditions (no reference counterpart: the reference runs one
 *      freestream condition per process, Config.cpp:479-488).  nbatch, a power
 *      of two in [1, 64], freestream conditions share the mesh and advance
 *      together, one launch per kernel for all of them; every per-cell host
 *      array of the state, time-step, gradient, limiter, residual and linear-
 *      solution calls below then holds nbatch blocks, condition-major
 *      ([nbatch][ncell][...]), and per-variable outputs hold [nbatch][nv].
 *      Changing the batch size discards the device state and the implicit
 *      workspace (call bes_implicit_setup again).  The implicit path takes a
 *      batch with GMRES and precond 0-2; each condition then runs its own
 *      Arnoldi process, stopping point and pseudo-time loop, and its results
 *      are the same bits as when it runs alone. ---- */
int bes_set_batch(bes_ctx *ctx, int nbatch);
int bes_get_batch(const bes_ctx *ctx);
/* limiter freeze (no reference counterpart; the reference recomputes the
 * Venkatakrishnan limiter at every pass, LimiterVenkatakrishnan.cpp:21-99).
 * drop in (0, 1): the implicit loop (bes_implicit_run/_continue) fixes a
 * condition's limiter values at those of the step whose density residual
 * first reaches drop times its first value; later gradient passes of that
 * condition (implicit and explicit) recompute the gradient but reuse those
 * limiter values.  A frozen limiter makes the residual a smooth function of
 * the state, which ends the limit cycles of a few limited cells that can hold
 * the residual of a fine mesh at a fixed level.  drop 0 (the default) never
 * freezes; any call clears the frozen state, and so does bes_set_batch.
 * bes_limiter_frozen: per condition 1 if frozen ([nbatch]). */
int bes_set_limiter_freeze(bes_ctx *ctx, double drop);
int bes_limiter_frozen(bes_ctx *ctx, int *frozen);
/* freeze every condition's limiter at given values, phi [nbatch][ncell][nv] as
 * bes_get_limiter returns them (no reference counterpart): later gradient
 * passes recompute the gradient and use these values, so that several solver
 * forms converge the same frozen discretisation.  bes_set_limiter_freeze
 * (any drop) ends it. */
int bes_freeze_limiter_values(bes_ctx *ctx, const double *phi);
/* freestream primitive per condition, [nbatch][nv+1] */
int bes_set_freestream_batch(bes_ctx *ctx, const double *primitive);

/* SolutionSnapshot::setConservedVariable / getConservedVariable for all cells
 * (SolutionSnapshot.cpp:60-143); U [nbatch][ncell][nv] */
int bes_set_state(bes_ctx *ctx, const double *U);
int bes_get_state(bes_ctx *ctx, double *U);

/* ---- TimeStepCalculator::computeAllTimeStep (TimeStepCalculator.hpp:43-46;
 *      TimeStepCalculatorLocal.cpp:67-93, TimeStepCalculatorGlobal.cpp:64-98) ---- */
int bes_compute_time_step(bes_ctx *ctx, double cfl, int local);
int bes_get_time_step(bes_ctx *ctx, double *dt /*[nbatch][ncell]*/);

/* ---- Gradient::computeGradient + Limiter::computeLimiterTerm (Gradient.hpp:44-49,
 *      Limiter.hpp:45-51; GradientGreenGauss.cpp:21-135,
 *      LimiterVenkatakrishnan.cpp:21-128): one fused pass over the current state ---- */
int bes_compute_gradient_limiter(bes_ctx *ctx);
int bes_get_gradient(bes_ctx *ctx, double *G /*[nbatch][ncell][nv][ndim]*/);
int bes_get_limiter(bes_ctx *ctx, double *phi /*[nbatch][ncell][nv]*/);

/* ---- second-order residual sum_f +-F*A per cell (the face loop of
 *      SolverExplicit.cpp:285-344 with Solver.cpp:69-121); uses the gradient and
 *      limiter of the last bes_compute_gradient_limiter call.  R_out
 *      [nbatch][ncell][nv] may be NULL; maxabs_out [nbatch][nv] (Linf per
 *      variable) may be NULL ---- */
int bes_compute_residual(bes_ctx *ctx, double *R_out, double *maxabs_out);

/* ---- SolverExplicit::timeStep_ (SolverExplicit.cpp:259-393): niter full
 *      multi-stage pseudo-time steps on the device state.  linf_out
 *      [niter][nbatch][nv] (last-stage Linf of the residual, the reference
 *      monitor) may be NULL. ---- */
int bes_explicit_run(bes_ctx *ctx, int niter, int nstage, const double *alpha, double cfl,
                     int local, double *linf_out);
/* on != 0: bes_explicit_run captures one pseudo-time step (time step, every
 * stage, the monitor) as a HIP graph and replays it niter times, one graph
 * launch per step (stage counts > 1, profiling off; no reference counterpart) */
int bes_explicit_graph(bes_ctx *ctx, int on);
/* on != 0: bes_explicit_run runs all niter steps as one cooperative launch of a
 * persistent kernel whose resident blocks sweep each pass (time step; per
 * stage gradient and limiter, face fluxes, residual and update) and meet at a
 * grid barrier between passes; results bitwise those of the launched passes.
 * Used for one condition with local time steps and profiling off, otherwise
 * the launched (or graph) form runs; takes precedence over the graph (no
 * reference counterpart; the reference's loop is SolverExplicit.cpp:395-438) */
int bes_explicit_persistent(bes_ctx *ctx, int on);

/* ---- Flux_::computeFlux / FluxJacobian::computeJacobian over a batch of faces
 *      (Flux.hpp:43-46, FluxJacobian.hpp:36-44; FluxRoe.tpp:306-333,
 *      FluxFreestream.tpp:192-226, FluxSlipWall.tpp:32-53, FluxJacobianWrapper.tpp:32-82).
 *      kind = BES_INTERNAL (Roe), BES_SLIP_WALL, BES_FREESTREAM.  UR rows have
 *      nv+1 entries: the neighbour conserved state (Roe) or the freestream
 *      primitive state (far field); unused otherwise.  Host arrays in and out. ---- */
int bes_flux_batch(int ndim, int kind, long n, const double *UL, const double *UR,
                   const double *normals, const bes_gas *gas, double *F);
int bes_jacobian_batch(int ndim, int kind, long n, const double *UL, const double *UR,
                       const double *normals, const bes_gas *gas, double *JL, double *JR);
/* bes_flux_batch without any allocation: no counterpart in the reference; for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results as
 * bes_flux_batch; inputs and outputs are staged through one fixed device array
 * per GPU (held until the process ends), so n is at most BES_NOALLOC_MAX_FACES
 * (BES_ERR_ARG above it).  Calls on one GPU share that array: calls that
 * overlap in time on one GPU (several host threads) race. */
enum { BES_NOALLOC_MAX_FACES = 16384 };
int bes_flux_batch_noalloc(int ndim, int kind, long n, const double *UL, const double *UR,
                           const double *normals, const bes_gas *gas, double *F);

/* ---- implicit path: SolverImplicit + CoefficientMatrix + the AMGCL solve ----
 * bes_implicit_setup: allocate the matrix and solver workspace.
 *   precision 64 (FP64 matrix, factors and Krylov vectors) or 32 (FP32 inner
 *   solve; the right-hand side and the update stay FP64); precond 0 none,
 *   1 block Jacobi, 2 block ILU(0) in multicolour order, 3 the learned
 *   operator (bes_lp_*, below; falls back to ILU(0) for a solve it stalls),
 *   4 aggregation AMG (bes_amg_*, below); GMRES(restart),
 *   restart <= 30, stops at ||b - A x|| <= rtol ||b|| (the reference's
 *   LINEAR_SOLVER_ABS_TOL is this relative tolerance: CoefficientMatrix.cpp:304).
 *   Replaces CoefficientMatrix::initialize (CoefficientMatrix.cpp:114-164). */
int bes_implicit_setup(bes_ctx *ctx, int precision, int precond, int restart, int maxiter, double rtol);
/* switch the preconditioner (0-4) of an existing setup; the next solve factors again */
int bes_implicit_set_precond(bes_ctx *ctx, int precond);
/* classical Gram-Schmidt re-orthogonalises (DGKS) when the new Krylov vector
 * keeps less than eta of its norm (default 0.7; 0 never) */
int bes_implicit_set_reorth(bes_ctx *ctx, double eta);
/* counters since setup: re-orthogonalisations, learned solves that fell back,
 * GMRES cycles whose Givens estimate met the tolerance while the true residual
 * did not (the solve then restarted from the true residual); any may be NULL */
int bes_implicit_stats(bes_ctx *ctx, long *reorths, int *lp_fallbacks, long *extra_restarts);
/* Krylov method of CoefficientMatrix::solve (LINEAR_SOLVER, Config.cpp:345-358;
 * CoefficientMatrix.cpp:282-296).  BES_SOLVER_GMRES: right-preconditioned
 * GMRES(restart); it also serves the reference's default FGMRES, which equals
 * right-preconditioned GMRES whenever the preconditioner is one fixed linear
 * operator during a solve, as every preconditioner here is.  BES_SOLVER_BICGSTAB:
 * AMGCL's BiCGStab with right preconditioning (bicgstab.hpp).  The reference's
 * LGMRES is not available: BES_ERR_ARG. */
enum { BES_SOLVER_GMRES = 0, BES_SOLVER_BICGSTAB = 1 };
int bes_implicit_set_solver(bes_ctx *ctx, int solver);
/* order of the multicolour block ILU(0) (no reference counterpart: AMGCL's
 * ILU0 is sequential): the colours come from a greedy pass over the cells in
 * device (Morton) order (0, the default), the mesh file's cell order (1), or
 * along the freestream velocity (2; uses the freestream set before).  Changes
 * the factors and so the iteration counts, never the solution of a solve that
 * converges.  bes_implicit_ncolors: colours of the current order. */
int bes_implicit_set_coloring(bes_ctx *ctx, int mode);
int bes_implicit_ncolors(bes_ctx *ctx);
/* time step, gradient, limiter, face Jacobians and the system A x = b at the
 * current state (SolverImplicit.cpp:307-320 = resetSystemOfEquation_ + dt +
 * gradient + limiter + computeFaceFluxes_).  repeat_nv != 0 reproduces the
 * reference's nV-fold addition of every Jacobian block (SolverImplicit.cpp:
 * 425-502).  linf_b [nv] = Linf of b (the reference's implicit monitor,
 * SolverImplicit.cpp:520-523), linf_r [nv] = Linf of sum_f +-F*A; may be NULL. */
int bes_implicit_assemble(bes_ctx *ctx, double cfl, int local, int repeat_nv, double *linf_b, double *linf_r);
/* the assembled system in the reference's CoefficientMatrix CSR layout
 * (CoefficientMatrix.cpp:127-160; MatrixSparse.hpp getRowPtr/getColPtr/getValPtr)
 * and right-hand side [ncell][nv]; nnz = bes_implicit_nnz(ctx) */
long bes_implicit_nnz(bes_ctx *ctx);
int bes_implicit_get_system(bes_ctx *ctx, long *rowptr, long *col, double *val, double *rhs);
/* load a system assembled by the caller in the same layout (the host-side
 * CoefficientMatrix::changeAllVal/addToDiagonal/addToVal/addToBlock path,
 * CoefficientMatrix.cpp:166-204); val [nnz], rhs [ncell][nv] */
int bes_implicit_set_system(bes_ctx *ctx, const double *val, const double *rhs);
/* CoefficientMatrix::solve (CoefficientMatrix.cpp:210-325): iterations and the
 * true relative residual ||b - A x|| / ||b||; BES_ERR_SOLVE if rtol was not met
 * (for a batch: the largest of the conditions'; per condition see
 * bes_implicit_batch_state) */
int bes_linear_solve(bes_ctx *ctx, int *iters, double *relres);
int bes_get_linear_solution(bes_ctx *ctx, double *x /*[ncell][nv]*/);
/* U += x (SolverImplicit::updateSolution_, SolverImplicit.cpp:509-528) */
int bes_implicit_update(bes_ctx *ctx);
/* per condition of the batch (arrays [nbatch], any may be NULL): steps taken
 * and drop reached in the loop of bes_implicit_run/_continue, the Courant
 * number of its next step, Krylov iterations and true relative residual of
 * the last solve (no reference counterpart) */
int bes_implicit_batch_state(bes_ctx *ctx, int *steps, int *reached, double *cfl, int *iters, double *relres);
/* SolverImplicit::solve (SolverImplicit.cpp:552-600): up to nsteps
 * backward-Euler steps with the adaptive Courant number of SolverImplicit.cpp:
 * 530-550 (cfl_min = COURANT_NUM, cfl_max = ADAPTIVE_COURANT_NUM_MAX,
 * exponents = ADAPTIVE_COURANT_NUM_EXPONENT).  Stops early when the Linf of
 * the density residual sum_f +-F*A falls below stop_drop times its first value
 * (stop_drop <= 0: never).  hist [nsteps][2*nv+3] (may be NULL): Linf(b) per
 * var, Linf(sum F*A) per var, Courant number, Krylov iterations, relres. */
int bes_implicit_run(bes_ctx *ctx, int nsteps, double cfl_min, double cfl_max, const double *exponents,
                     int local, int repeat_nv, double stop_drop, double *hist, int *steps_done);
/* With a batch, hist is [nsteps][nbatch][2*nv+3], every condition keeps its
 * own Courant number and stops (is frozen: no solve, no update) once its drop
 * is reached; the rows of frozen conditions are NaN; linf_b and linf_r of
 * bes_implicit_assemble are [nbatch][nv].
 * nsteps more steps of the loop the last bes_implicit_run started (same
 * parameters; Courant number and residual history carry over), stopping early
 * at its stop_drop; reached = 1 once the drop has been reached. */
int bes_implicit_continue(bes_ctx *ctx, int nsteps, double *hist, int *steps_done, int *reached);

/* ---- learned preconditioner (no reference counterpart; an alternative to the
 *      AMGCL preconditioner of CoefficientMatrix.cpp:210-325, used with
 *      precond 3).  A message-passing network on the cell dual graph (width 32,
 *      3 layers, BF16 matrix-core kernels with FP32 accumulation) picks, per
 *      face and row, a blend G in (0,1) between block Jacobi and the
 *      first-order Neumann approximation of A^-1:
 *        (M v)_i = D_i^-1 v_i - sum_m G_im D_i^-1 O_im D_j^-1 v_j.
 *      It is trained on the GPU on the assembled system by Adam on the loss
 *      mean_q ||r_K||^2/||v_q||^2, r_{k+1} = r_k - A M r_k, r_0 = v_q (probe
 *      q = 0 the right-hand side, the rest seeded Gaussian vectors).
 * bes_lp_init: network state after bes_implicit_setup; deterministic weights
 *   from seed, zero decoder weights with bias 3 (G = 0.953 everywhere).
 * bes_lp_set_params / bes_lp_get_params: the flat FP32 parameter vector
 *   (n = bes_lp_param_count()).
 * bes_lp_operator: builds M for the assembled system; Md [ncell][nv][nv],
 *   Me [ncell][k][nv][nv] (by local face), G [ncell][k][nv]; any may be NULL.
 * bes_lp_loss_grad: loss and gradient for caller-given probe vectors
 *   [nprobe][ncell][nv] (reference cell order), nrich Richardson steps.
 * bes_lp_train: steps Adam steps (learning rate lr) on the assembled system;
 *   loss_hist [steps] may be NULL.
 * bes_lp_fallbacks: solves so far that fell back to ILU(0). ---- */
int bes_lp_param_count(void);
int bes_lp_init(bes_ctx *ctx, unsigned long long seed);
int bes_lp_set_params(bes_ctx *ctx, const float *params, int n);
int bes_lp_get_params(bes_ctx *ctx, float *params, int n);
int bes_lp_operator(bes_ctx *ctx, double *Md, double *Me, double *G);
int bes_lp_loss_grad(bes_ctx *ctx, const double *probes, int nprobe, int nrich, double *loss, float *grad);
int bes_lp_train(bes_ctx *ctx, int steps, double lr, int nprobe, int nrich, unsigned long long seed,
                 double *loss_hist);
int bes_lp_fallbacks(bes_ctx *ctx);
/* online training (no reference counterpart): with steps > 0, every solve of
 * the learned form first runs that many Adam steps (learning rate lr, the
 * right-hand side and nprobe - 1 seeded Gaussian probes, nrich Richardson
 * steps) on the system the implicit loop has just assembled, so the network is
 * trained on the Jacobians and residuals the solver itself generates; the
 * operator is then rebuilt for the solve.  steps 0 switches it off.
 * bes_lp_online_loss: the loss of the last online step. */
int bes_lp_set_online(bes_ctx *ctx, int steps, double lr, int nprobe, int nrich);
double bes_lp_online_loss(bes_ctx *ctx);
/* the operator is rebuilt at every n-th solve after an assembly and reused by
 * the solves between (1: every solve); new parameters or training force a
 * rebuild.  bes_lp_builds: operator builds so far (no reference counterpart) */
int bes_lp_set_rebuild_interval(bes_ctx *ctx, int n);
long bes_lp_builds(bes_ctx *ctx);

/* ---- aggregation AMG preconditioner (precond 4): the AMGCL hierarchy of
 *      CoefficientMatrix.cpp:210-325 with coarsening "aggregation" and
 *      aggr.block_size = nv (:231-256), rebuilt on the GPU for the solve after
 *      an assembly (or every rebuild_interval-th).  FP32 solves sweep row-scaled
 *      FP16 copies of the level matrices.  Strength eps_strong on the largest |entry| of each block
 *      (AMGCL's plain_aggregates / pointwise_aggregates); aggregates of radius
 *      two around roots that are at least three strong links apart, chosen in
 *      parallel rounds (the reference sweeps the rows in order); coarse
 *      operators P^T A P / over_interp; levels while the scalar rows exceed
 *      coarse_enough, the coarsest inverted densely; V-cycle from x = 0 with one
 *      pre- and one post-sweep (amg.hpp:287-304, 498-541).
 * bes_amg_set_params: eps_strong (0.08), over_interp (1.5), damping (0.72),
 *   coarse_enough (500 scalar rows), relax 0 damped Jacobi on the scalar
 *   diagonal (the reference's DAMPED_JACOBI), 1 SPAI0 (its SPAI0), 2 damped
 *   block Jacobi on the nv x nv diagonal blocks (no reference counterpart);
 *   rebuild_interval: the hierarchy is rebuilt at every rebuild_interval-th
 *   solve and reused by the solves between (1: every solve, as AMGCL is).
 * bes_amg_info: levels, block rows and stored blocks per level (up to maxlev),
 *   direct = 1 when the coarsest level is inverted, builds so far.
 * bes_amg_level: the aggregate (next-level row, -1 removed) and state (1 root,
 *   2 member, 3 removed) of every row of a level (level 0 in the reference's
 *   cell order), and for levels >= 1 the block CSR of the level's matrix
 *   (rowptr [n+1], col [nnzb] ascending, val [nnzb][nv][nv]); any may be NULL.
 * bes_amg_apply: one V-cycle z = M r on host vectors [ncell][nv] (reference
 *   cell order), building the hierarchy for the assembled system if needed. ---- */
int bes_amg_set_params(bes_ctx *ctx, double eps_strong, double over_interp, double damping, int coarse_enough,
                       int relax, int rebuild_interval);
/* aggregation rule: 0 parallel rounds (default), 1 AMGCL's own row sweep in
 * the reference's cell order (plain_aggregates.hpp:137-206), run serially by
 * one GPU thread: the reference's aggregates exactly, for verification-size
 * systems */
int bes_amg_set_aggregation(bes_ctx *ctx, int mode);
int bes_amg_info(bes_ctx *ctx, int *nlevels, long *rows, long *nnzb, int *direct, long *builds, int maxlev);
int bes_amg_level(bes_ctx *ctx, int level, long *agg, unsigned char *state, long *rowptr, long *col, double *val);
int bes_amg_apply(bes_ctx *ctx, const double *r, double *z);

/* ---- solution files in the reference's formats, condition cond of the batch:
 *      format 0 legacy VTK (SolutionSnapshot.cpp:172-266), 1 CSV of centroid and
 *      primitives (:146-170); the slip-wall CSV of face centre, area vector and
 *      owner primitives (:268-308).  Primitives are written in the reference's
 *      dimensional form (GasPerfect.tpp:294-320). ---- */
int bes_write_solution(bes_ctx *ctx, const char *path, int format, int cond);
int bes_write_slip_wall(bes_ctx *ctx, const char *path, int cond);

/* ---- measurement (no counterpart in the reference, which times nothing):
 *      when profiling is on, every kernel launch is bracketed by HIP events on
 *      the context's stream; bes_get_profile returns, per kernel kind, the
 *      summed device time, the launch count and the summed algorithmic HBM
 *      bytes (every array a launch must touch, counted once; neighbour
 *      gathers served by cache not counted) since profiling was switched on.
 *      kinds: 0 time step, 1 gradient+limiter, 2 residual+update,
 *      3 Jacobian, 4 assembly, 5 SpMV, 6 preconditioner apply, 7 Krylov vector
 *      ops, 8 factorisation, 9 implicit update, 10 face flux, 11 learned
 *      preconditioner (build and training) ---- */
int bes_set_profiling(bes_ctx *ctx, int on);
/* autotuner (no reference counterpart; the reference's counterpart is its
 * OpenMP schedule, SolverExplicit.cpp:285): times each tunable kernel kind at
 * block sizes 64, 128 and 256 on the context's current data with HIP events
 * and keeps the fastest.  flags: 1 the explicit kernels (gradient and
 * limiter, time step, face flux, residual) on the current state; 2 the implicit kernels (Jacobians, assembly,
 * factorisation, SpMV, preconditioner, update) on the assembled system, which
 * is assembled and factored again (same values); 4 the ILU(0) colouring order
 * (bes_implicit_set_coloring), chosen as the one whose solve of the assembled
 * system is fastest (the linear solution is overwritten).  chosen
 * [16] (block size per kernel kind, as bes_get_profile numbers them), best_us
 * [16] (their time, 0 when not tuned) and coloring may be NULL. */
int bes_autotune(bes_ctx *ctx, int flags, int *chosen, double *best_us, int *coloring);
int bes_get_profile(bes_ctx *ctx, double *total_ms, long *launches, double *bytes, int nkinds);
int bes_synchronize(bes_ctx *ctx);

/* ---- error reporting ---- */
long bes_last_error_face(const bes_ctx *ctx);      /* reference face id of the first NaN flux */
const char *bes_last_error_message(void);
const char *bes_status_string(int status);

#ifdef __cplusplus
}
#endif
#endif /* BES_HIP_H */
