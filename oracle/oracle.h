/*
 * oracle.h -- CPU restatement of the reference's generic 3-D ADER-DG kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against, and the CPU baseline that bench.py
 * times beside it.  Nothing in exahype_amd/ may include, link or call it.
 *
 * Every function restates one reference function, with the same loop
 * order and the same floating-point expression trees (built with
 * -ffp-contract=off, like g++'s x86-64 default), cited at its definition
 * in oracle.c.  All arrays use the reference layouts (row-major, last index
 * fastest):
 *   luh    (z,y,x,var)               B^3 * (V+P)
 *   lduh   (z,y,x,var)               B^3 * V
 *   lQi    (z,y,x,t,var)             B^4 * (V+P)
 *   rhs    (t,z,y,x,var)             B^4 * V
 *   lFi    (t,z,y,x,{x,y,z,src},var) B^4 * 4V
 *   gradQ  (z,y,x,t,dir,var)         B^4 * 3V
 *   lQhi   (z,y,x,var)               B^3 * (V+P)
 *   lFhi   4 blocks of B^3*V: x (z,y,x), y (z,x,y), z (y,x,z), src (z,y,x)
 *   lQhbnd (face,a,b,var) 6*B^2*(V+P); lFhbnd (face,a,b,var) 6*B^2*V
 *          x-faces (z,y), y-faces (z,x), z-faces (y,x)
 */
#ifndef EXAHYPE_ORACLE_H
#define EXAHYPE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORA_PDE_EULER = 0, ORA_PDE_ELASTIC = 1 };

typedef struct ora_solver {
  int pde;              /* ORA_PDE_* */
  int order;            /* N; basis size B = N+1, 0 <= N <= 9 */
  int nvar;             /* V */
  int npar;             /* P (material parameters stored after the variables) */
  int use_flux;
  int use_source;
  int use_ncp;
  int no_time_averaging;
  double cfl;
} ora_solver;

/* ---- PDE callbacks (user solver interface restated) ---- */
void ora_flux(const ora_solver* s, const double* Q, double** F);
void ora_eigenvalues(const ora_solver* s, const double* Q, int d, double* lambda);
void ora_algebraic_source(const ora_solver* s, const double* Q, double* S);
void ora_ncp(const ora_solver* s, const double* Q, const double* gradQ, double* BgradQ);

/* ---- per-cell kernels ---- */
/* Returns the number of Picard iterations performed (1 .. 2B). */
int ora_space_time_predictor_nonlinear(const ora_solver* s, double* lQhbnd, double* lFhbnd,
                                       double* lQi, double* rhs, double* lFi, double* gradQ,
                                       double* lQhi, double* lFhi,
                                       const double* luh, const double* dx, double dt);
void ora_volume_integral_nonlinear(const ora_solver* s, double* lduh, const double* lFi,
                                   const double* lFhi, const double* dx);
void ora_surface_integral_nonlinear(const ora_solver* s, double* lduh, const double* lFhbnd,
                                    const double* dx);
void ora_riemann_solver_nonlinear(const ora_solver* s, double* FL, double* FR, const double* QL,
                                  const double* QR, double dt, int direction);
void ora_solution_update(const ora_solver* s, double* luh, const double* lduh, double dt);
double ora_stable_time_step_size(const ora_solver* s, const double* luh, const double* dx);

/* table access (the same reference tables the kernels use) */
double ora_table(int which, int order, int i, int j);
enum { ORA_T_WEIGHTS = 0, ORA_T_NODES, ORA_T_FLCOEFF, ORA_T_FRCOEFF, ORA_T_KXI, ORA_T_IK1,
       ORA_T_DUDX, ORA_T_LOB_WEIGHTS, ORA_T_LOB_NODES };

/* ---- uniform periodic grid driver (non-fused global time stepping) ---- */
typedef struct ora_grid ora_grid;
/* nx*ny*nz cells on [lo,hi], cell c = (k*ny + j)*nx + i. */
ora_grid* ora_grid_create(const ora_solver* s, int nx, int ny, int nz, const double* lo,
                          const double* hi);
void ora_grid_destroy(ora_grid* g);
long ora_grid_num_cells(const ora_grid* g);
void ora_grid_set_solution(ora_grid* g, const double* luh);
void ora_grid_get_solution(const ora_grid* g, double* luh);
/* dt = min over cells of the stable time step of the current solution. */
double ora_grid_stable_dt(ora_grid* g, int nthreads);
/* Advance nsteps time steps; dt_out[nsteps] (may be NULL) receives the dt of
 * each step; iters_out[nsteps*ncells] (may be NULL) the Picard counts. */
void ora_grid_step(ora_grid* g, int nsteps, int nthreads, double* dt_out, int* iters_out);
double ora_grid_time(const ora_grid* g);
/* Picard counts of the most recent predictor pass, [ncells]. */
void ora_grid_picard_iterations(const ora_grid* g, int* out);
/* Timing sample: all per-cell stages (predictor, volume, Riemann on the
 * cell's three upper faces paired within the sample, surface, update,
 * dt) for the first m cells; solution of those cells is advanced. */
void ora_grid_sample_step(ora_grid* g, long m, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
