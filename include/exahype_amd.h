/*
 * exahype_amd.h -- C-ABI of the MI355X-native ADER-DG time-stepping core.
 *
 * Plain pointers and sizes only.  Array layouts are those of the reference
 * (row-major, variable innermost), B = order+1, D = nvar+npar:
 *   luh    (z,y,x,var)               B^3 D        lduh   (z,y,x,var)  B^3 V
 *   lQhbnd (face,a,b,var)            6 B^2 D      lFhbnd (face,a,b,var) 6 B^2 V
 *          faces -x,+x,-y,+y,-z,+z; x-faces (z,y), y-faces (z,x), z-faces (y,x)
 *   lQi    (z,y,x,t,var)             B^4 D        lFi (t,z,y,x,{x,y,z,src},var) 4 B^4 V
 *   lQhi   (z,y,x,var)               B^3 D        lFhi 4 blocks of B^3 V:
 *          x (z,y,x), y (z,x,y), z (y,x,z), src (z,y,x)
 * Every function returns an exa_status.  Host pointers unless stated.
 *
 * Reference interfaces replaced (paths relative to the reference checkout,
 * K = ExaHyPE/kernels/aderdg/generic):
 *   kernels::aderdg::generic::c::*   K/Kernels.h:85-213 and the K/c/3d/ sources
 *   the per-cell loop of ADERDGSolver (ExaHyPE/exahype/solvers/ADERDGSolver.cpp:
 *   1881 performPredictionAndVolumeIntegral, 2185 updateSolution, 2789
 *   solveRiemannProblemAtInterface, 2004 computeTimeStepSize) driven by the
 *   Runner's non-fused step (ExaHyPE/exahype/runners/Runner.cpp:1116-1162).
 */
#ifndef EXAHYPE_AMD_H
#define EXAHYPE_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum exa_status {
  EXA_OK = 0,
  EXA_E_HIP = 1,         /* a HIP runtime call failed */
  EXA_E_UNSUPPORTED = 2, /* no compiled instantiation for this solver */
  EXA_E_INVALID = 3,     /* bad argument */
  EXA_E_NOMEM = 4,       /* device allocation failed */
  EXA_E_NONFINITE = 5    /* a non-finite time step size was produced */
} exa_status;

/* PDEs compiled into the library (the user's SolverType becomes a device functor). */
typedef enum exa_pde { EXA_PDE_EULER = 0, EXA_PDE_ELASTIC = 1 } exa_pde;

/* The spec file's `terms` (useFlux / useSource / useNCP template flags). */
enum { EXA_TERM_FLUX = 1, EXA_TERM_SOURCE = 2, EXA_TERM_NCP = 4 };

/* Arithmetic of the bulk work.  EXA_PRECISION_FP64: the reference's double
 * everywhere.  EXA_PRECISION_FP32_BULK: space-time predictor, volume integral
 * and Riemann solver in float on float face / update arrays; surface integral,
 * solution update and the time step size in double on the double solution.
 * The float Picard loop stops on a relative residual (RMS change < 1e-6 RMS
 * of the iterate) because the reference's absolute 1e-14 lies below float
 * rounding.  Grid entry points only; the single-cell ones are FP64. */
enum { EXA_PRECISION_FP64 = 0, EXA_PRECISION_FP32_BULK = 1 };

/* Space-time predictor variants (the device kernel behind
 * spaceTimePredictorNonlinear + volumeIntegralNonlinear).  AUTO: the default,
 * the tensor contractions on the vector ALUs with a 144 KB LDS budget for the
 * Picard iterate (measured fastest); VALU_LDS64 / VALU_LDS144: the vector-ALU
 * kernel with a 64 KB / 144 KB LDS budget; MFMA: the FP32-bulk form of the
 * Euler solvers (no parameters, orders 2..7) with the contractions on the
 * matrix cores (v_mfma_f32_4x4x1f32).  The variants compute the same
 * mathematics and differ in rounding order only. */
enum { EXA_PREDICTOR_AUTO = 0, EXA_PREDICTOR_VALU_LDS64 = 1, EXA_PREDICTOR_VALU_LDS144 = 2, EXA_PREDICTOR_MFMA = 3 };

/* SolverType's compile-time constants (AbstractGenericADERDGSolverHeader.template:32-35). */
typedef struct exa_solver {
  int pde;               /* exa_pde */
  int order;             /* SolverType::Order (N) */
  int num_variables;     /* SolverType::NumberOfVariables (V) */
  int num_parameters;    /* SolverType::NumberOfParameters (P) */
  int terms;             /* EXA_TERM_* */
  int no_time_averaging; /* template flag noTimeAveraging */
  double cfl;            /* SolverType::CFL */
  int precision;         /* EXA_PRECISION_* (0 = FP64) */
  int linear;            /* 1: a linear ADER-DG solver -- the generated glue calls the *Linear
                            kernels (AbstractGenericADERDGSolverImplementation.template:53-86) */
} exa_solver;

const char* exa_status_string(int status);
/* Selects the HIP device for the calling thread; optional (device 0 by default). */
int exa_init(int device);
/* 1 if (pde, order, nvar, npar, terms) has a compiled instantiation. */
int exa_solver_supported(const exa_solver* s);

/* ---------------------------------------------------------------------------
 * Single-cell entry points: the kernels::aderdg::generic::c call signatures
 * (host arrays in, host arrays out; synchronous).  The scratch arguments of
 * the reference become optional outputs of its temporaries (NULL = skip).
 * ------------------------------------------------------------------------- */

/* spaceTimePredictorNonlinear (K/c/3d/spaceTimePredictorNonlinear.cpph:661).
 * lQi = tempSpaceTimeUnknowns[0], lFi = tempSpaceTimeFluxUnknowns[0],
 * lQhi = tempUnknowns, lFhi = tempFluxUnknowns.  picard_iterations may be NULL. */
int exa_space_time_predictor_nonlinear(const exa_solver* s, double* lQhbnd, double* lFhbnd,
                                       double* lQi, double* lFi, double* lQhi, double* lFhi,
                                       const double* luh, const double* dx, double dt,
                                       int* picard_iterations);

/* volumeIntegralNonlinear (K/c/3d/volumeIntegralNonlinear.cpph:31).  lFhi is the
 * time-averaged flux (4 B^3 V); with s->no_time_averaging the reference passes the
 * space-time fluxes lFi (4 B^4 V) instead, and so does this argument. */
int exa_volume_integral_nonlinear(const exa_solver* s, double* lduh, const double* lFhi,
                                  const double* dx);

/* surfaceIntegralNonlinear (K/c/3d/surfaceIntegralNonlinear.cpph:21); lduh in/out. */
int exa_surface_integral_nonlinear(const exa_solver* s, double* lduh, const double* lFhbnd,
                                   const double* dx);

/* riemannSolverNonlinear (K/c/3d/riemannSolverNonlinear.cpph:37); FL, FR in/out. */
int exa_riemann_solver_nonlinear(const exa_solver* s, double* FL, double* FR, const double* QL,
                                 const double* QR, double dt, int direction);

/* ---- linear ADER-DG kernels (FP64; the entry point implies the solver's linear flag).
 * Supported: elasticity (9 variables, 3 parameters) at orders 2..5 and the
 * five-variable Euler test solver with one parameter at order 3. */

/* spaceTimePredictorLinear (K/c/3d/spaceTimePredictorLinear.cpph:34), the
 * Cauchy-Kowalevski predictor.  lQi, lFi, gradQ, lQhi, lFhi are the reference's
 * temporaries (tempSpaceTimeUnknowns[0], tempSpaceTimeFluxUnknowns[0..1],
 * tempUnknowns, tempFluxUnknowns), written when all five are given (or all
 * null): lQi (t,z,y,x,var) (B+1) B^3 (V+P); lFi (t,z,y,x,dir,var) 3 B^4 V plus
 * a zeroed source block of B^4 V; gradQ (t,z,y,x,dir,var) 3 B^4 V; lQhi B^3 (V+P);
 * lFhi (dir,z,y,x,var) 3 B^3 V plus a zeroed source block of B^3 V.  PSi:
 * point sources (t,z,y,x,var) (B+1) B^3 V (tempPointForceSources), or null. */
int exa_space_time_predictor_linear(const exa_solver* s, double* lQhbnd, double* lFhbnd, double* lQi,
                                    double* lFi, double* gradQ, double* lQhi, double* lFhi,
                                    const double* luh, const double* dx, double dt, const double* PSi);

/* The projection step of pointSource (K/c/3d/pointSource.cpph:17-126): on entry
 * pointForceSources holds the user's pointSource values at the (B+1) B^3
 * space-time nodes ((t,z,y,x,var); time node 0 at t, node n >= 1 at t + dt
 * x_{n-1}; space node at center + dx (x_node - 1/2)) and x0 the source position
 * that call returned (3 doubles per node); on return each value is the point
 * force projected onto the cell's nodal basis, or zero where x0 lies outside
 * the cell.  Orders 1..7.  kernels.hpp's pointSource<SolverType> makes the
 * user calls on the host and then calls this. */
int exa_point_source_projection(const exa_solver* s, double* pointForceSources, const double* x0,
                                const double* center, const double* dx);
/* Gauss-Legendre node i of [0,1] at an order (the table the kernels use). */
double exa_quadrature_node(int order, int i);

/* volumeIntegralLinear (K/c/3d/volumeIntegralLinear.cpph:22); lFhi 4 B^3 V
 * (the source block is not read); dx is unused, as in the reference. */
int exa_volume_integral_linear(const exa_solver* s, double* lduh, const double* lFhi, const double* dx);

/* surfaceIntegralLinear (K/c/3d/surfaceIntegralLinear.cpph:27); lduh in/out. */
int exa_surface_integral_linear(const exa_solver* s, double* lduh, const double* lFhbnd,
                                const double* dx);

/* riemannSolverLinear (K/c/3d/riemannSolverLinear.cpph:25); FL, FR out
 * (fluctuations); dt is unused, as in the reference. */
int exa_riemann_solver_linear(const exa_solver* s, double* FL, double* FR, const double* QL,
                              const double* QR, double dt, int direction);

/* ---- output: the Cartesian VTK plotter's vertex values (FP64) ----
 * ADERDG2CartesianVTK::plotVertexData (exahype/plotters/VTK/ADERDG2CartesianVTK.cpp:288):
 * each cell's V+P values at its (N+1)^3 equidistant vertices offset + i h/N
 * (equidistantGridProjector1d), [cells][(N+1)^3][V+P] with x fastest; the
 * identity post-processing.  Orders 1..6.  exahype_amd/vtk.py writes them as a
 * legacy VTK file. */
int exa_plot_vertices(const exa_solver* s, long ncells, const double* luh, double* vertices);

/* ---- AMR transfer operators (FP64; K/c/3d/amrRoutines.cpph, Kernels.h:163-199).
 * Blocks: volumes luh (z,y,x,var) B^3 (V+P); faces lQhbnd (b,a,var) B^2 (V+P) and
 * lFhbnd B^2 V of one face.  subfaceIndex[2] / subcellIndex[3] in [0, 3^(fine -
 * coarse level)) per direction (faces: the face's two in-plane directions in
 * storage order; volumes: x, y, z).  Prolongation overwrites the fine block
 * (fine = coarse level: no change, as in the reference); restriction adds the
 * restricted fine block to the coarse block and needs fine > coarse level.
 * Orders 1..5; level differences up to 6. */

/* faceUnknownsProlongation (amrRoutines.cpph:48). */
int exa_face_unknowns_prolongation(const exa_solver* s, double* lQhbndFine, double* lFhbndFine,
                                   const double* lQhbndCoarse, const double* lFhbndCoarse,
                                   int coarseGridLevel, int fineGridLevel, const int* subfaceIndex);
/* faceUnknownsRestriction (amrRoutines.cpph:183). */
int exa_face_unknowns_restriction(const exa_solver* s, double* lQhbndCoarse, double* lFhbndCoarse,
                                  const double* lQhbndFine, const double* lFhbndFine,
                                  int coarseGridLevel, int fineGridLevel, const int* subfaceIndex);
/* volumeUnknownsProlongation (amrRoutines.cpph:313). */
int exa_volume_unknowns_prolongation(const exa_solver* s, double* luhFine, const double* luhCoarse,
                                     int coarseGridLevel, int fineGridLevel, const int* subcellIndex);
/* volumeUnknownsRestriction (amrRoutines.cpph:420). */
int exa_volume_unknowns_restriction(const exa_solver* s, double* luhCoarse, const double* luhFine,
                                    int coarseGridLevel, int fineGridLevel, const int* subcellIndex);

/* Device buffers for the batched entry points (hipMalloc / hipFree / copies). */
void* exa_device_alloc(size_t bytes);
int exa_device_free(void* p);
int exa_copy_to_device(void* dst, const void* src, size_t bytes);
int exa_copy_to_host(void* dst, const void* src, size_t bytes);

/* Batched transfers on device arrays of blocks (all pointers from
 * exa_device_alloc; no counterpart in the reference, which calls the kernels
 * above once per cell or face pair from its grid traversal).  items: device
 * array of 5 ints per transfer: destination block, source block, subinterval
 * index per direction (faces use two; face blocks are numbered cell * 6 + face
 * in lQhbnd/lFhbnd layouts).  Prolongation: nitems transfers, each destination
 * overwritten.  Restriction: ndst destinations; ptr (device, ndst + 1 ints)
 * gives each destination's items [ptr[k], ptr[k+1]) of the nitems, all with
 * that destination block; their restricted blocks are added to it in list
 * order, as successive single calls would.  Each destination block may appear
 * once per call (in one prolongation item, or in one restriction range); a
 * repeated destination is written by one of its transfers only, the others are
 * skipped, and the call returns EXA_E_INVALID.  The calls block until done;
 * items with an index out of range are skipped and the call returns
 * EXA_E_INVALID.
 * elapsed_ms (optional): the kernels' time from HIP events on their stream. */
int exa_volume_unknowns_prolongation_batch(const exa_solver* s, double* luhFine, long nfine,
                                           const double* luhCoarse, long ncoarse, const int* items,
                                           long nitems, int levelDelta, float* elapsed_ms);
int exa_volume_unknowns_restriction_batch(const exa_solver* s, double* luhCoarse, long ncoarse,
                                          const double* luhFine, long nfine, const int* ptr, long ndst,
                                          const int* items, long nitems, int levelDelta, float* elapsed_ms);
int exa_face_unknowns_prolongation_batch(const exa_solver* s, double* lQhbndFine, double* lFhbndFine,
                                         long nfine_faces, const double* lQhbndCoarse,
                                         const double* lFhbndCoarse, long ncoarse_faces, const int* items,
                                         long nitems, int levelDelta, float* elapsed_ms);
int exa_face_unknowns_restriction_batch(const exa_solver* s, double* lQhbndCoarse, double* lFhbndCoarse,
                                        long ncoarse_faces, const double* lQhbndFine, const double* lFhbndFine,
                                        long nfine_faces, const int* ptr, long ndst, const int* items,
                                        long nitems, int levelDelta, float* elapsed_ms);

/* exa_volume_unknowns_restriction_batch without any device allocation.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, semantics and results as the batch call
 * (without its timing argument); the error flag and the destination claims
 * live in a fixed device array of the library, one per GPU for the life of
 * the process, resolved by exa_init (or the first call on a GPU): at most
 * 8,388,608 destination blocks (ncoarse) per call, beyond which the call
 * returns EXA_E_INVALID before launching anything.  Every call on a GPU
 * reuses that GPU's array, so calls that overlap on one GPU (several host
 * threads, or several processes sharing it) race: serialise them.
 * Synchronous. */
int exa_volume_unknowns_restriction_batch_noalloc(const exa_solver* s, double* luhCoarse, long ncoarse,
                                                  const double* luhFine, long nfine, const int* ptr, long ndst,
                                                  const int* items, long nitems, int levelDelta);

/* solutionUpdate (K/c/3d/solutionUpdate.cpph:23); luh in/out. */
int exa_solution_update(const exa_solver* s, double* luh, const double* lduh, double dt);

/* stableTimeStepSize (K/c/3d/stableTimeStepSize.cpph:21). */
int exa_stable_time_step_size(const exa_solver* s, const double* luh, const double* dx,
                              double* dt_out);

/* ---------------------------------------------------------------------------
 * Batched entry points: the same kernels over `ncells` cells in one launch
 * (host arrays, cells contiguous; dx is [ncells][3]).
 * ------------------------------------------------------------------------- */

/* Predictor + volume integral (performPredictionAndVolumeIntegral,
 * ADERDGSolver.cpp:1881).  lduh may be NULL; picard_iterations[ncells] may be NULL. */
int exa_batch_predictor_volume(const exa_solver* s, long ncells, double* lQhbnd, double* lFhbnd,
                               double* lduh, const double* luh, const double* dx, double dt,
                               int* picard_iterations);

/* ---------------------------------------------------------------------------
 * Device-resident grid: one grid level as an array of cells with host-built
 * face lists; the non-fused global time step of the reference
 * (Riemann on every face once -> surface integral + update -> dt = min over
 * cells -> predictor + volume integral) runs entirely on the device.
 * ------------------------------------------------------------------------- */
typedef struct exa_grid exa_grid;

/* EXA_BC_USER: the PDE's boundary_values (boundaryValues,
 * UserSolverInterface.h:171) at every node of a boundary face, the outside
 * state left of even faces and right of odd ones (ADERDGSolver.cpp:2952-2960);
 * for PDEs that provide the hook (user PDEs, exahype_amd/user_solver.hpp). */
enum { EXA_BC_PERIODIC = 0, EXA_BC_TRANSMISSIVE = 1, EXA_BC_USER = 2 };

/* Uniform nx*ny*nz Cartesian grid on [lo, hi]; cell c = (k*ny + j)*nx + i;
 * boundary EXA_BC_PERIODIC or EXA_BC_TRANSMISSIVE (boundaryConditions,
 * ExaHyPE/kernels/aderdg/generic/c/3d/boundaryConditions.cpph:24, with
 * boundaryValues returning the inside state).  A solver with linear = 1 runs
 * the linear kernels (Cauchy-Kowalevski predictor + volumeIntegralLinear,
 * riemannSolverLinear, surfaceIntegralLinear + solutionUpdate); at a
 * transmissive boundary face its fluctuations vanish (outside = inside). */
int exa_grid_create(const exa_solver* s, int nx, int ny, int nz, const double* lo,
                    const double* hi, int boundary, exa_grid** out);
int exa_grid_destroy(exa_grid* g);
long exa_grid_num_cells(const exa_grid* g);
int exa_grid_upload_solution(exa_grid* g, const double* luh);
int exa_grid_download_solution(exa_grid* g, double* luh);
/* Solution of cells [first, first+count). */
int exa_grid_download_cells(exa_grid* g, long first, long count, double* luh);

/* Euler initial data: rho = rho0 + amp sin(2 pi k.x) + sum of nmodes seeded
 * Fourier modes (amplitude mode_amp, integer wave vectors in [-3,3]^3,
 * phases in [0, 2 pi)) in rho and u,v,w; velocity vel; constant pressure p. */
typedef struct exa_euler_ic {
  double rho0, amp;
  double k[3];
  double vel[3];
  double p;
  int nmodes;
  double mode_amp;
  unsigned long long seed;
} exa_euler_ic;
int exa_grid_init_euler(exa_grid* g, const exa_euler_ic* ic);
/* Elastic initial data (EXA_PDE_ELASTIC, 9 variables + lambda, mu, rho): zero
 * stresses, a Gaussian velocity pulse u = v = w = amp exp(-|x-c|^2/(2 width^2))
 * at the centre of [0,1]^3, and a seeded two-layer, smoothly modulated
 * material with rho in [2,3], cp in [4,6], cs = cp/sqrt(3). */
int exa_grid_init_elastic(exa_grid* g, unsigned long long seed, double width, double amp);

/* Advance nsteps time steps (asynchronous on the grid's stream). */
int exa_grid_step(exa_grid* g, int nsteps);

/* The grid's vertex values for the Cartesian VTK plotter (exa_plot_vertices'
 * layout, cells in grid order), computed on the device and copied to the host;
 * elapsed_ms (optional): the kernel's time from HIP events on the grid's stream. */
int exa_grid_plot_vertices(exa_grid* g, double* vertices, float* elapsed_ms);
int exa_grid_synchronize(exa_grid* g);
/* Current time stamp, time step size and number of completed steps. */
int exa_grid_time(exa_grid* g, double* t, double* dt, long* steps);
/* dt of steps 0..n-1 (dt used by step k = out[k]). */
int exa_grid_dt_history(exa_grid* g, double* out, long n);
/* Picard iteration counts of the most recent predictor launch, [ncells]. */
int exa_grid_picard_iterations(exa_grid* g, int* out);

/* Per-stage timing with HIP events recorded on the grid's stream. */
enum { EXA_STAGE_PREDICTOR = 0, EXA_STAGE_RIEMANN = 1, EXA_STAGE_UPDATE = 2, EXA_STAGE_DT = 3,
       EXA_NUM_STAGES = 4 };
int exa_grid_set_profiling(exa_grid* g, int on);
/* Predictor variant of the grid's later predictor launches (EXA_PREDICTOR_*;
 * default AUTO).  EXA_E_UNSUPPORTED for a variant the solver does not compile. */
int exa_grid_set_predictor_variant(exa_grid* g, int variant);
/* Corrector of the nonlinear solvers' steps.  on (the default): one kernel per
 * step solves the Riemann problems of each cell's six faces from the cell's
 * and its neighbours' face blocks and applies surface integral, solution
 * update and the stable dt (the fluxes of a face are identical in both cells);
 * off: a face pass writing both cells' flux blocks, then the update pass.
 * Linear solvers always run the two passes (EXA_E_UNSUPPORTED for on). */
int exa_grid_set_fused_corrector(exa_grid* g, int on);
/* Fused time stepping: the optimisation `fuse-algorithmic-steps` with
 * `fuse-algorithmic-steps-factor` weight in (0, 1] (Parser.cpp:486-505;
 * Runner::runOneTimeStepWithFusedAlgorithmicSteps, Runner.cpp:1053-1110).
 * Each step corrects with the corrector dt and predicts with the predictor
 * dt chosen one step earlier; the new solution's stable dt s then checks it
 * (TimeStepSizeComputation.cpp:167-204): when it exceeded s, the prediction
 * is re-run with dt = weight*s, which the next step also corrects with;
 * otherwise the next predictor dt is (used + weight*s)/2.  The first step
 * uses the stable dt of the current solution (no weighting: the reference
 * weighs the first dt only in MPI builds, TimeStepSizeComputation.cpp:
 * 229-234).  Takes effect at the next exa_grid_step, which starts with a new
 * dt and prediction.  Nonlinear solvers only (EXA_E_UNSUPPORTED for linear);
 * exa_grid_time reports the corrector dt, exa_grid_dt_history the corrector
 * dt of every step. */
int exa_grid_set_fused_time_stepping(exa_grid* g, int on, double weight);
/* Predictions re-run after a failed stability check (Runner.cpp:773). */
int exa_grid_predictor_reruns(exa_grid* g, long* reruns);

/* ---------------------------------------------------------------------------
 * User PDEs on the device.  The reference's user solver implements flux,
 * eigenvalues, nonConservativeProduct, algebraicSource, boundaryValues and
 * adjustPointSolution (UserSolverInterface.h:64-171) and the generic kernels
 * are templates over it.  Here a user writes those as a struct of static
 * __device__ functions and instantiates the library's kernel templates for it
 * in a shared object of their own with exahype_amd/user_solver.hpp
 * (EXA_USER_SOLVER), which exports this table; the grid below drives it.
 * ------------------------------------------------------------------------- */
#define EXA_USER_SOLVER_ABI 1
typedef struct exa_user_solver {
  int abi_version;      /* EXA_USER_SOLVER_ABI */
  int order;            /* SolverType::Order */
  int num_variables;    /* SolverType::NumberOfVariables */
  int num_parameters;   /* SolverType::NumberOfParameters */
  int terms;            /* EXA_TERM_* */
  double cfl;           /* SolverType::CFL */
  int has_boundary_values;  /* boundaryValues compiled (EXA_BC_USER available) */
  int has_adjust;           /* adjustPointSolution compiled */
  /* Kernel launches on a HIP stream (passed as void*); the argument structs
   * are the library's (include/exahype_amd/device/aderdg.hpp); the return
   * value is a hipError_t (0: launched). */
  int (*predictor)(const void* predictor_args, long ncells, int precision, void* stream);
  int (*corrector)(const void* corrector_args, int precision, void* stream);
  int (*stable_dt)(const void* update_args, long ncells, void* stream);
  int (*adjust)(const void* adjust_args, long ncells, void* stream);
} exa_user_solver;

/* A grid (as exa_grid_create) whose PDE is a user solver; precision
 * EXA_PRECISION_*; boundary EXA_BC_PERIODIC, EXA_BC_TRANSMISSIVE or, when the
 * solver has boundary values, EXA_BC_USER.  The steps run the one-kernel
 * corrector.  Initial data: exa_grid_upload_solution, or
 * exa_grid_adjust_solution (adjustPointSolution at t = 0, as
 * setInitialConditions does, ADERDGSolver.cpp:2161-2180). */
int exa_grid_create_user(const exa_user_solver* us, int precision, int nx, int ny, int nz, const double* lo,
                         const double* hi, int boundary, exa_grid** out);
/* Apply the PDE's adjustPointSolution at every node after each update, at
 * t + dt (ADERDGSolver.cpp:2231-2236); EXA_E_UNSUPPORTED without the hook. */
int exa_grid_set_adjust_solution(exa_grid* g, int on);
/* solutionAdjustment of every cell now, at the grid's time with dt = 0
 * (solutionAdjustment.cpph:17-47); EXA_E_UNSUPPORTED without the hook. */
int exa_grid_adjust_solution(exa_grid* g);
/* Sum of per-cell Picard iteration counts and of cells over every predictor
 * launch made while profiling was on (the work of the predictor launches). */
int exa_grid_picard_totals(exa_grid* g, unsigned long long* iterations, unsigned long long* cells);
/* Synchronises, then returns per-stage summed milliseconds and launch counts
 * since profiling was switched on (and resets them). */
int exa_grid_stage_times(exa_grid* g, double* ms, long* launches);

/* ---------------------------------------------------------------------------
 * A-posteriori subcell limiter (kernels::limiter::generic::c and the
 * MUSCL-Hancock finite-volume kernels).  Euler, FP64, orders 2..3; FV patches
 * are (L+2g)^3 cells with L = 2N+1, variable fastest, host arrays.
 * ------------------------------------------------------------------------- */

/* projectOnFVLimiterSpace (ExaHyPE/kernels/limiter/generic/c/Limiter.cpp:23):
 * writes the interior of lim[(L+2g)^3 * V] from luh[B^3 * V]; ghosts untouched. */
int exa_project_on_fv_limiter_space(const double* luh, int num_variables, int basis_size,
                                    int ghost_layer_width, double* lim);
/* projectOnDGSpace (Limiter.cpp:69): luh[B^3 * V] from the interior of lim. */
int exa_project_on_dg_space(const double* lim, int num_variables, int basis_size,
                            int ghost_layer_width, double* luh);
/* findCellLocalMinAndMax (Limiter.cpp:210), observables = the variables:
 * extrema over Gauss-Legendre nodes, Gauss-Lobatto nodes and subcell means. */
int exa_find_cell_local_min_and_max(const double* luh, int num_variables, int basis_size,
                                    double* local_min, double* local_max);
/* discreteMaximumPrincipleAndMinAndMaxSearch (Limiter.cpp:114): bmin/bmax hold
 * 6 face blocks of num_variables; faces 0/1 are overwritten with the local /
 * boundary extrema; *dmp_ok = 1 when the DMP holds. */
int exa_discrete_maximum_principle_and_min_and_max_search(const double* luh, int num_variables,
                                                          int basis_size, double relaxation,
                                                          double difference_scaling, double* bmin,
                                                          double* bmax, int* dmp_ok);
/* Finite-volume schemes (the spec file's FV `type` / limiter-type). */
enum { EXA_FV_GODUNOV = 0, EXA_FV_MUSCLHANCOCK = 1 };

/* musclhancock::c::solutionUpdate (ExaHyPE/kernels/finitevolumes/musclhancock/
 * c/3d/musclhancock.cpph:49) for Euler: luh_new = update(luh) on one patch
 * with filled ghost layers (patch_size L, ghost_layer_width 2); dx is the
 * DG cell size.  Only the interior of luh_new is written.  *dt_max_allowed
 * (may be NULL) receives the function's return value: the minimum over the
 * patch's faces of 0.99/3 * h_d / s_max. */
int exa_musclhancock_solution_update(const exa_solver* s, double* luh_new, const double* luh,
                                     const double* dx, double dt, int patch_size,
                                     int ghost_layer_width, double* dt_max_allowed);

/* godunov::c::solutionUpdate (ExaHyPE/kernels/finitevolumes/godunov/c/3d/
 * godunov.cpph:35) for Euler (flux only): first-order update of one patch
 * (L+2g)^3 with filled ghost layers, any patch_size L >= 1 and
 * ghost_layer_width g >= 1 (the reference's Godunov solvers use g = 1).
 * luh_new must hold a copy of luh on entry (FiniteVolumesSolver.cpp:669-671)
 * and receives the reference's writes: the interior update and the
 * boundary-face fluxes added to the first ghost layer.  dx is the patch
 * (cell) size.  *dt_max_allowed (may be NULL) receives the function's return
 * value: the minimum over the patch's faces of s->cfl/3 * h_d / s_max. */
int exa_godunov_solution_update(const exa_solver* s, double* luh_new, const double* luh,
                                const double* dx, double dt, int patch_size,
                                int ghost_layer_width, double* dt_max_allowed);

/* ---------------------------------------------------------------------------
 * Device-resident finite-volumes grid: one uniform level of
 * FiniteVolumesSolver cells (ExaHyPE/exahype/solvers/FiniteVolumesSolver.cpp),
 * each a patch of L^3 subcells, under global time stepping.  One step:
 * neighbour merging (ghostLayerFilling, FVS:767-803; transmissive boundary
 * FVS:805-846), updateSolution with the scheme's solutionUpdate (FVS:662-715),
 * startNewTimeStep's stableTimeStepSize of the new solution with the solver's
 * CFL (FVS:553-574), dt = minimum over cells.  Euler (flux only), FP64.
 * EXA_FV_GODUNOV: subcells live in dense planes without ghost layers:
 * neighbour reads wrap (periodic) or clamp to the boundary subcell
 * (transmissive, the reference's mirror for ghost width 1), which gives the
 * values the reference's ghost layers hold.  EXA_FV_MUSCLHANCOCK: every cell
 * keeps its patch with two ghost layers (its results depend on the patches:
 * the face-ghost subcells' transverse slopes read the never-filled, zero edge
 * ghosts, as in the reference); patch sizes 5 and 7.
 * ------------------------------------------------------------------------- */
typedef struct exa_fvgrid exa_fvgrid;
/* nx*ny*nz cells on [lo, hi], cell c = (k*ny + j)*nx + i, patch_size L;
 * boundary EXA_BC_PERIODIC or EXA_BC_TRANSMISSIVE (boundaryValues returning
 * the inside state).  EXA_E_UNSUPPORTED for MUSCL-Hancock with L other than
 * 5 or 7. */
int exa_fvgrid_create(const exa_solver* s, int scheme, int patch_size, int nx, int ny, int nz,
                      const double* lo, const double* hi, int boundary, exa_fvgrid** out);
int exa_fvgrid_destroy(exa_fvgrid* g);
long exa_fvgrid_num_cells(const exa_fvgrid* g);
/* Solution [cells][L^3][5] (subcells z,y,x, variable fastest).  Upload and
 * init_euler restart the clock: t = 0 and dt = the new solution's stable dt. */
int exa_fvgrid_upload(exa_fvgrid* g, const double* interior);
int exa_fvgrid_download(exa_fvgrid* g, double* interior);
/* The exa_grid_init_euler state evaluated at the subcell centres. */
int exa_fvgrid_init_euler(exa_fvgrid* g, const exa_euler_ic* ic);
/* Advance nsteps time steps (asynchronous on the grid's stream). */
int exa_fvgrid_step(exa_fvgrid* g, int nsteps);
int exa_fvgrid_synchronize(exa_fvgrid* g);
/* Time stamp, time step size of the next step and completed steps;
 * EXA_E_NONFINITE once a step produced a non-finite or non-positive dt. */
int exa_fvgrid_time(exa_fvgrid* g, double* t, double* dt, long* steps);
/* HIP-event timing of the update kernel (Godunov + stable dt of the new
 * solution) and of the time-stamp kernel per step while profiling is on;
 * stage_times synchronises, returns the sums and launch counts, and resets. */
enum { EXA_FV_STAGE_UPDATE = 0, EXA_FV_STAGE_FINALIZE = 1, EXA_FV_NUM_STAGES = 2 };
int exa_fvgrid_set_profiling(exa_fvgrid* g, int on);
int exa_fvgrid_stage_times(exa_fvgrid* g, double* ms, long* launches);

/* Device-resident limiting grid: LimitingADERDGSolver's non-fused time step
 * (ExaHyPE/exahype/solvers/LimitingADERDGSolver.cpp:806-1742, Runner.cpp:866-946,
 * 1116-1162) -- DG / FV neighbour merging by limiter status, status-dependent
 * update, DMP + admissibility, status spreading, rollback and local
 * recomputation -- on one uniform level. */
typedef struct exa_lgrid exa_lgrid;
/* boundary: EXA_BC_PERIODIC or EXA_BC_TRANSMISSIVE; helper_layers h gives
 * troubled status 2h+1 and FV-active status >= h+1 (ADERDGSolver.cpp:461-462).
 * fv_scheme: the spec file's limiter-type, EXA_FV_MUSCLHANCOCK or
 * EXA_FV_GODUNOV (the generated FV solver's solutionUpdate,
 * Toolkit/src/eu/exahype/solvers/GenericFiniteVolumes{MUSCLHancock,Godunov}InC.java). */
int exa_lgrid_create(const exa_solver* s, int nx, int ny, int nz, const double* lo, const double* hi,
                     int boundary, int helper_layers, int steps_till_cured, double dmp_relaxation,
                     double dmp_difference_scaling, int fv_scheme, exa_lgrid** out);
int exa_lgrid_destroy(exa_lgrid* g);
long exa_lgrid_num_cells(const exa_lgrid* g);
/* Initial state: DG solution luh [cells][B^3][V] and the FV initial condition
 * at subcell centres lim_ic [cells][L^3][V] (may be NULL: projection of luh),
 * used for the patches of initially troubled cells and their helpers. */
int exa_lgrid_upload(exa_lgrid* g, const double* luh, const double* lim_ic);
/* Spherical blast initial condition generated on the device: rho = 1 +
 * noise (2u - 1), u from splitmix64 keyed by the global node (subcell) index,
 * zero velocity, p = p_in inside |x - centre| < radius else p_out. */
int exa_lgrid_init_blast(exa_lgrid* g, unsigned long long seed, double radius, double p_in,
                         double p_out, double noise);
int exa_lgrid_step(exa_lgrid* g, int nsteps);
/* Fused time stepping of the limiting solver (fuse-algorithmic-steps with
 * weight in (0, 1]; LimitingADERDGSolver::fusedTimeStep, LS:806-836, in
 * Runner::runOneTimeStepWithFusedAlgorithmicSteps, Runner.cpp:1053-1110).
 * Each step merges, updates and sets the limiter status as the non-fused
 * step, then predicts the cells below troubled with the predictor dt and
 * checks that dt as exa_grid_set_fused_time_stepping does.  After an
 * irregular limiter-domain change the spreading, rollback and local
 * recomputation of the non-fused step follow (updateMeshAndSubdomains,
 * Runner.cpp:866-946); the cells whose solution came from an FV patch are
 * predicted again with the step's predictor dt (recomputePredictorLocally,
 * LS:1438-1464), the check is repeated with the recomputed solution's stable
 * dt s (LocalRecomputation.cpp:168-222), and corrector and predictor dt both
 * become s (updateTimeStepSizesFused of the adapter's TimeStepSizeComputation,
 * ADERDGSolver.cpp:594-612).  A failed last check re-runs the prediction with
 * the new dt.  Switching on takes the current dt as the predictor dt: set
 * before the first step (after upload or init) this is the fused start, the
 * unweighted stable dt of the initial state. */
int exa_lgrid_set_fused_time_stepping(exa_lgrid* g, int on, double weight);
/* Predictions re-run after a failed stability check. */
int exa_lgrid_predictor_reruns(exa_lgrid* g, long* reruns);
int exa_lgrid_synchronize(exa_lgrid* g);
int exa_lgrid_time(exa_lgrid* g, double* t, double* dt, long* steps);
int exa_lgrid_download(exa_lgrid* g, double* luh);
/* Limiter status per cell, and the number of steps with an irregular change. */
int exa_lgrid_status(exa_lgrid* g, int* status, long* irregular_steps);
/* Current FV patch of one cell, (L+4)^3 * V.  The interior is the cell's FV
 * solution; the ghost layers are scratch (refilled from the neighbours before
 * every update).  The reference's update also adds the boundary-face fluxes
 * to the first ghost layer of its new patch (musclhancock.cpph:267-329), values
 * nothing reads before the next refill; the device does not write them. */
int exa_lgrid_limiter_solution(exa_lgrid* g, long cell, double* lim);
/* Solution min/max face blocks [cells][6][V]. */
int exa_lgrid_min_max(exa_lgrid* g, double* mn, double* mx);
/* Stage timing: EXA_STAGE_PREDICTOR, _RIEMANN, _UPDATE (the status kernel:
 * DG surface integral and update, DMP, admissibility and status), _DT,
 * EXA_STAGE_LIMITER (merging of min/max and ghost layers, spreading,
 * rollback, recomputation) and EXA_STAGE_FV (the FV update of the
 * FV-active cells' patches). */
enum { EXA_STAGE_LIMITER = 4, EXA_STAGE_FV = 5, EXA_NUM_LSTAGES = 6 };
int exa_lgrid_set_profiling(exa_lgrid* g, int on);
int exa_lgrid_stage_times(exa_lgrid* g, double* ms, long* launches);
/* Picard iterations summed over the predictor launches since profiling was
 * switched on, and the cells those launches covered (FV-active cells, which
 * the predictor skips, count 0 iterations). */
int exa_lgrid_picard_totals(exa_lgrid* g, unsigned long long* iterations, unsigned long long* cells);

/* ---------------------------------------------------------------------------
 * Patch set: many independent periodic grids ("patches") of uneven size, order
 * and PDE advanced together, each with its own time step size.  Every patch
 * is the unit cube with n^3 cells: an Euler patch starts from the entropy wave
 * rho = 1 + amp sin(2 pi k.x), u = v = w = 1, p = 1; an elastic patch from the
 * seeded material and velocity pulse of exa_grid_init_elastic.  Each of them
 * follows the non-fused global time step of exa_grid
 * (ADERDGSolver::updateSolution, performPredictionAndVolumeIntegral,
 * solveRiemannProblemAtInterface, ADERDGSolver.cpp:1881-2290; Runner.cpp:
 * 1116-1162) with its own dt.  Patches of one PDE and order share every
 * launch (cell -> patch map); the patches run in waves whose state fits
 * hbm_budget_gb; each wave's time step is captured once in a hipGraph and
 * replayed.  The reference has no batching of
 * independent grids: this scheduler replaces running one ExaHyPE process per
 * patch.
 * ------------------------------------------------------------------------- */
typedef struct exa_patch_spec {
  int n;       /* cells per side */
  int order;   /* polynomial order: Euler 2..7, elastic 2..5 */
  double k[3]; /* Euler: wave vector of the entropy wave */
  double amp;  /* Euler: density amplitude of the entropy wave; elastic: pulse amplitude */
  int pde;     /* EXA_PDE_EULER (0) or EXA_PDE_ELASTIC: velocity-stress elasticity,
                  seeded two-layer material and a Gaussian velocity pulse
                  (as exa_grid_init_elastic) */
  unsigned long long seed; /* elastic: material seed */
  double width;            /* elastic: pulse width */
} exa_patch_spec;
typedef struct exa_patchset exa_patchset;
/* precision: EXA_PRECISION_*; the solution of every patch with at most
 * keep_max_cells cells is copied to host memory at the end of its wave (for
 * exa_patchset_solution; 0 keeps none).  EXA_E_NOMEM when one patch alone
 * exceeds the budget. */
int exa_patchset_create(const exa_patch_spec* patches, int npatches, int precision, double hbm_budget_gb,
                        long keep_max_cells, exa_patchset** out);
/* Set up every wave (initial condition, first dt and predictor), optionally
 * autotune the predictor variant per order (EXA_PREDICTOR_*), and advance each patch
 * warmup_steps + nsteps time steps, replaying a captured hipGraph when
 * use_graph != 0; the last nsteps are timed.  Synchronous. */
int exa_patchset_run(exa_patchset* b, int warmup_steps, int nsteps, int use_graph, int autotune_variants);
/* Device seconds of the timed steps (HIP events around each wave's timed steps),
 * degrees of freedom updated in them, and the number of waves. */
int exa_patchset_stats(exa_patchset* b, double* seconds, double* dof_updates, int* nwaves);
/* Predictor variant the autotuner chose for the Euler patches of an order
 * (EXA_PREDICTOR_*; -1 none); exa_patchset_group_variant: for the patches of
 * one PDE (EXA_PDE_*) and order. */
int exa_patchset_predictor_variant(exa_patchset* b, int order);
int exa_patchset_group_variant(exa_patchset* b, int pde, int order);
/* Corrector of the patch set's steps: on (the default), the one-kernel
 * corrector (each cell solves its six faces' Riemann problems inside its
 * patch, then surface integral, update and dt); off, the face pass and the
 * update pass.  Applies to later exa_patchset_run calls. */
int exa_patchset_set_fused_corrector(exa_patchset* b, int on);
/* Profiling of later exa_patchset_run calls: the timed steps run eagerly
 * (no graph) with HIP events around every corrector and predictor launch and
 * the Picard iterations counted per group; exa_patchset_group_times returns
 * one group's totals over the timed steps of all waves (predictor and
 * corrector ms, predictor launches, Picard iterations, predicted cells). */
int exa_patchset_set_profiling(exa_patchset* b, int on);
int exa_patchset_group_times(exa_patchset* b, int pde, int order, double* predictor_ms, double* corrector_ms,
                             long* launches, unsigned long long* picard_iterations,
                             unsigned long long* predictor_cells);
/* Solution [n^3][B^3][V+P] of one patch after the run (kept patches only). */
int exa_patchset_solution(exa_patchset* b, int patch, double* luh);
/* Time stamp and the dt of the next step of one patch after the run. */
int exa_patchset_time(exa_patchset* b, int patch, double* t, double* dt);
int exa_patchset_destroy(exa_patchset* b);

#ifdef __cplusplus
}
#endif
#endif
