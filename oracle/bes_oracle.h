/* CPU oracle for the BES residual / implicit-update path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain CPU restatement of the reference
 * algorithm (joel-ho/BES), used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  The product path never links,
 * loads or calls it.
 *
 * Conventions: conserved state is cell-major [ncell][nv] (the reference's
 * SolutionSnapshot layout), gradients [ncell][nv][nd], face arrays in the
 * reference face order (boundary faces in file order, then internal faces in
 * discovery order, Mesh.cpp:236-534).
 *
 * geometry mode bits: 0 = reference behaviour ("compat");
 *   1 = true tetrahedron volume (undo quirk: MeshGmsh.cpp:101-133 omits /3),
 *   2 = true tetrahedron centroid (undo quirk: MeshGmsh.cpp:135-145 counts v2 twice),
 *   4 = per-face reconstruction vector (undo quirk: Element.cpp:102-105 keys its
 *       face map before face ids exist, so Solver.cpp:91 reads local face 0).
 */
#ifndef BES_ORACLE_H
#define BES_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_mesh or_mesh;

/* gas / scheme constants: gamma, gas constant R, Harten limit, Venkatakrishnan K */
typedef struct { double gamma, R, harten, venkat_k; } or_gas;

or_mesh *or_mesh_create(int ndim, long nnode, const double *coords, long ncell,
                        const long *cells, long nbface, const long *bfaces,
                        const int *btype, int geom_mode);
void or_mesh_destroy(or_mesh *m);
long or_mesh_nface(const or_mesh *m);
int or_mesh_kface(const or_mesh *m);
/* any output pointer may be NULL */
void or_mesh_get_faces(const or_mesh *m, long *owner, long *nbr, int *btype,
                       double *S, double *n, double *area, double *center);
void or_mesh_get_cells(const or_mesh *m, double *vol, double *centroid, long *cface,
                       long *cadj, double *w, double *r, long *rsel, double *sproj);

/* per-face fluxes, kind 1 = Roe, 2 = slip wall, 3 = freestream.  UR holds the
 * neighbour conserved state (Roe) or the freestream primitive state [nv+1]
 * (freestream); record stride of UR is nv+1. */
void or_flux(int ndim, int kind, long n, const double *UL, const double *UR,
             const double *nrm, const or_gas *gas, double *F);
/* forward-mode AD Jacobians with the reference's (Eigen 3.3.4) derivative
 * rules; JL = dF/dUL, JR = dF/dUR (Roe only; may be NULL), row-major nv x nv */
void or_jacobian(int ndim, int kind, long n, const double *UL, const double *UR,
                 const double *nrm, const or_gas *gas, double *JL, double *JR);

void or_gradient(const or_mesh *m, const double *U, double *G);
void or_limiter(const or_mesh *m, const double *U, const double *G, const or_gas *gas,
                double *phi);
/* local != 0: per-cell dt; else the global minimum broadcast to every cell */
void or_timestep(const or_mesh *m, const double *U, double cfl, int local,
                 const or_gas *gas, double *dt);
/* second-order residual sum_f (+/-) F*A per cell; F (per face) may be NULL */
void or_residual(const or_mesh *m, const double *U, const double *G, const double *phi,
                 const double *finf, const or_gas *gas, double *R, double *F);
/* niter multi-stage explicit iterations in place on U (SolverExplicit.cpp:259-393);
 * linf [niter][nv] = Linf of the last-stage residual; nthreads <= 0 -> default */
void or_explicit(const or_mesh *m, double *U, int niter, int nstage, const double *alpha,
                 double cfl, int local, const double *finf, const or_gas *gas,
                 double *linf, int nthreads);
/* implicit system (SolverImplicit.cpp:347-507): BSR blocks in the reference
 * column order (own block, then neighbours in local face order).
 * browptr [ncell+1], bcol [nnzb], bval [nnzb][nv][nv], rhs [ncell][nv]. */
long or_implicit_nnzb(const or_mesh *m);
void or_implicit_assemble(const or_mesh *m, const double *U, double cfl, int local,
                          const double *finf, const or_gas *gas, int repeat_nv,
                          long *browptr, long *bcol, double *bval, double *rhs,
                          double *dt_out);

/* implicit pseudo-time loop (SolverImplicit.cpp:552-600) with OpenMP: face-
 * parallel assembly with atomic adds (SolverImplicit.cpp:358-507), then
 * block ILU(0) in a multicolour order + right-preconditioned GMRES(restart)
 * to ||b - A x|| <= rtol ||b|| (the reference's AMGCL solve cannot be built
 * here), U += x, adaptive Courant number (SolverImplicit.cpp:530-550).
 * nthreads <= 0: the OpenMP default. */
typedef struct or_implicit or_implicit;
or_implicit *or_implicit_create(const or_mesh *m, int restart, int maxiter, double rtol, int nthreads);
void or_implicit_destroy(or_implicit *s);
int or_implicit_ncolor(const or_implicit *s);
/* up to nsteps steps in place on U.  *cfl: Courant number of the next step
 * (updated); lb_old [nv]: Linf of b of the previous step (zeros at the start of
 * a run: the first step's rule then clamps to cfl_min); stops after the step
 * whose density Linf of sum F*A is <= stop_drop * r0 (r0 <= 0: this call's
 * first step).  hist [nsteps][2nv+3]: Linf(b) per var, Linf(sum F*A) per var,
 * Courant number, Krylov iterations, true relative residual.  Returns steps done. */
int or_implicit_steps(or_implicit *s, const or_mesh *m, double *U, int nsteps, double *cfl,
                      double cfl_min, double cfl_max, const double *ex, int local, int repeat_nv,
                      double stop_drop, double r0, const double *finf, const or_gas *gas,
                      double *lb_old, double *hist);
/* the last step's matrix (BSR values as or_implicit_assemble), right-hand side
 * and linear solution; any may be NULL */
void or_implicit_last(const or_implicit *s, double *bval, double *rhs, double *x);
/* limiter values [ncell][nv] that every later assembly uses instead of
 * recomputing the limiter (the product's limiter freeze; no reference
 * counterpart), or NULL to recompute again */
void or_implicit_freeze_limiter(or_implicit *s, const double *phi);

#ifdef __cplusplus
}
#endif
#endif
