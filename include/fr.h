/*
 * fr.h -- C ABI of libfrhip, the MI355X (gfx950) flux-reconstruction residual
 * engine.  Plain C types only: pointers, sizes, ints and doubles.
 *
 * The ABI replaces the hot path of FlurryPP's `solver` class (reference
 * include/solver.hpp:25-289, src/solver.cpp).  Each entry point names the
 * reference interface it stands in for.
 *
 * Conventions
 *  - every call returns 0 on success or a negative FR_E* code; the message of
 *    the last failure on the calling thread is in fr_last_error().  The
 *    reference calls fatalError() -> exit(1) (include/error.hpp:10-16) for the
 *    same conditions; this library never exits;
 *  - an fr_ctx owns all device memory; host buffers belong to the caller and
 *    are copied during the call;
 *  - host arrays passed with layout FR_LAYOUT_REF use the reference layout
 *    (src/matrix.cpp:260-272, last index fastest): U(spt, ele, field) at
 *    U[field + nFields*(ele + nEles*spt)];
 *  - device work is ordered on the context's HIP stream; getters synchronize;
 *  - one context per host thread; a context is not thread-safe.
 */
#ifndef FR_H
#define FR_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fr_ctx fr_ctx;

enum {
  FR_OK = 0,
  FR_EINVAL = -1,   /* bad argument or unsupported configuration */
  FR_EMESH = -2,    /* mesh / connectivity error (geo.cpp fatalError cases) */
  FR_EJACOBIAN = -3,/* negative Jacobian (solver.cpp:1283-1284) */
  FR_EHIP = -4,     /* HIP runtime error or missing device */
  FR_ENAN = -5,     /* NaN in residual (output.cpp:833-839) */
  FR_ENOMEM = -6,
  FR_EIO = -7       /* file cannot be opened, read or written (solution files, restart) */
};

/* FR_LAYOUT_DEVICE: the library's own [ele][field][pt] layout with the
 * elements in device order (fr_get_element_order) */
enum { FR_LAYOUT_REF = 0, FR_LAYOUT_DEVICE = 1 };

/* boundary-condition ids: include/global.hpp:63-79 */
enum {
  FR_BC_CHAR = 1, FR_BC_SUP_IN = 2, FR_BC_SUP_OUT = 3, FR_BC_SUB_IN = 4,
  FR_BC_SUB_OUT = 5, FR_BC_SLIP_WALL = 8, FR_BC_ISOTHERMAL_NOSLIP = 9,
  FR_BC_ADIABATIC_NOSLIP = 10, FR_BC_SYMMETRY = 12,
  FR_BC_ISOTHERMAL_NOSLIP_MOVING = 13
};

enum { FR_PREC_FP64 = 0, FR_PREC_FP32 = 1 };

/* Parameters of the path; mirrors the `input` fields it reads
 * (include/input.hpp, defaults src/input.cpp:214-481). */
typedef struct fr_config {
  int equation;    /* 0 advection-diffusion, 1 Euler / Navier-Stokes */
  int nDims;       /* 2 (quads) or 3 (hexes) */
  int order;       /* P, 1..5 */
  int riemannType; /* 0 Rusanov, 1 Roe (2D) */
  int viscous;
  int timeType;    /* 0 FE, 2 RK2, 4 RK44 ("RK45"), 5, 6 */
  int sptsType;    /* 0 Gauss-Legendre, 1 Gauss-Lobatto */
  int vcjhScheme;  /* 0 DG, 1 SD, 2 HU, 3 C+ */
  int fixVis;
  int viscFix;     /* 0 reproduce reference viscous index defects, 1 corrected */
  double dt;
  double gamma, advectVx, advectVy, advectVz, lambda, diffD;
  double rhoBound, uBound, vBound, wBound, pBound;
  double penFact, tau, mu_inf, prandtl, rt_inf, c_sth;
  double TWall, RGas, uWall, vWall, wWall;
} fr_config;

/* Unstructured mesh.  Hexes of 8 nodes (linear), 20 (quadratic
 * serendipity) or 27, 64, 125, 216 (Lagrange, equispaced, of order 2..5)
 * in the node order geo::readGmsh stores (geo.cpp:597-647: Gmsh's order,
 * the 20-node edges permuted) and shape_hex uses (funcs.cpp:178-566); quads
 * of 4 nodes; triangles of 3.  Faces are matched by their corner nodes.
 * vmap (optional) gives a canonical vertex id used only for face matching:
 * identifying the vertices of opposite box faces makes them periodic. */
typedef struct fr_mesh {
  int nVerts;
  const double *xv;   /* [nVerts][3] */
  int nEles;
  int nNodes;         /* nodes per element: 8, 20, 27, 64, 125, 216 (hex), 4 (quad), 3 (triangle) */
  const int *c2v;     /* [nEles][nNodes] */
  const int *vmap;    /* [nVerts] or NULL */
  int nBounds;
  const int *bcType;  /* [nBounds] */
  const int *bndOff;  /* [nBounds+1] CSR offsets into bndPts */
  const int *bndPts;  /* vertex ids on each boundary */
} fr_mesh;

/* solver::setup(input*, order, geo*) -- solver.cpp:21-63: builds the FR
 * operators (operators.cpp:13-76), face connectivity (geo.cpp:75-446,
 * intFace.cpp:6-71) and metrics (solver.cpp:109-210, 1227-1356) on device. */
int fr_create(const fr_config *cfg, const fr_mesh *mesh, int precision, fr_ctx **out);
int fr_destroy(fr_ctx *ctx);
const char *fr_last_error(void);

/* sizes: out[0..7] = nSpts, nFpts, nEles, nFields, nDims, nFaces, nIntFaces, nRKSteps */
int fr_get_sizes(fr_ctx *ctx, int *out8);

/* oper::opp_* public matrices (operators.hpp:133-146) rebuilt on the host:
 * which = 0 spts_to_fpts, 1 grad_spts[D], 2 extrapolateFn[D], 3 correction,
 * 4 correctU[D] (dense, row-major, reference shapes) */
int fr_get_operator(fr_ctx *ctx, int which, double *out);

/* geometry in reference layout: 0 JGinv_spts(dim,spt,ele,dim),
 * 1 detJac_spts(spt,ele), 2 norm_fpts(fpt,ele,dim), 3 dA_fpts(fpt,ele),
 * 4 pos_spts(spt,ele,dim) (solver.hpp geometry arrays, solver.cpp:109-210) */
int fr_get_geometry(fr_ctx *ctx, int which, double *out);

/* Device element order: order[i] = mesh element stored at device position i.
 * The library groups face-connected elements into the blocks its kernels
 * share (no reference counterpart; geo keeps the mesh order). */
int fr_get_element_order(fr_ctx *ctx, int *order /* nEles */);

/* face table as built by geo::setupElesFaces (geo.cpp:322-446): per face
 * (L ele, L local face, R ele, R local face, relRot, bcType or -1) */
int fr_get_faces(fr_ctx *ctx, int *out6);

/* U_spts setter / getter (solver.hpp:55; solver::initializeSolution writes it) */
int fr_set_state(fr_ctx *ctx, const double *U, int layout);
int fr_get_state(fr_ctx *ctx, double *U, int layout);

/* solver::calcResidual(step) -- solver.cpp:252-291: divF_spts[step] from the
 * current U_spts */
int fr_calc_residual(fr_ctx *ctx, int step);
/* divF_spts[step] (solver.hpp:63) */
int fr_get_residual(fr_ctx *ctx, int step, double *out, int layout);
/* U_fpts (solver.hpp:55): the face traces of the current state */
int fr_get_traces(fr_ctx *ctx, double *out /* U_fpts(fpt,ele,field) */);
/* The residual's intermediate arrays, the data members solver.hpp:57-60 that
 * the ele accessors read (ele.cpp:68-97): a debug path.  The fused kernels
 * keep them in registers and LDS; this call evaluates the residual of the
 * current state once (as fr_calc_residual(ctx, 0): divF_spts[0] and the traces
 * are those of the current state afterwards) with debug outputs switched on,
 * in buffers it allocates and frees.  which, in the reference layout:
 *   2 F_spts(dim,spt,ele,field)    the transformed total flux JGinv.F
 *   3 Fn_fpts(fpt,ele,field)       the common normal flux times dA, per side
 *   4 disFn_fpts(fpt,ele,field)    after correctDivFlux: Fn_fpts - disFn
 *   5 dU_spts(dim,spt,ele,field)   viscous: the corrected physical gradient
 *   6 dU_fpts(dim,fpt,ele,field)   viscous: the gradient at the flux points
 * FP64 quad and hex contexts; FR_EINVAL otherwise, and for 5 / 6 on the
 * inviscid path. */
int fr_get_array(fr_ctx *ctx, int which, double *out);

/* solver::update() -- solver.cpp:212-250, repeated nSteps times */
int fr_update(fr_ctx *ctx, int nSteps);
/* solver::timeStepA / timeStepB / copyUspts_U0 / copyU0_Uspts
 * (solver.cpp:307-399) */
int fr_time_step_a(fr_ctx *ctx, int step, double rkval);
int fr_time_step_b(fr_ctx *ctx, int step);
int fr_copy_u_u0(fr_ctx *ctx);
int fr_copy_u0_u(fr_ctx *ctx);

/* reductions: kind 1,2,3 = residual norm of divF_spts[0] as ele::getNormResidual
 * + writeResidual (ele.cpp:1882-1903, output.cpp:846-892), out[nFields];
 * kind 10 = kinetic energy integral of 0.5*rho*|u|^2 (out[0]) and volume
 * (out[1]) by the spt quadrature of solver.cpp:1510-1526 */
int fr_reduce(fr_ctx *ctx, int kind, double *out);

/* fr_reduce without device allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Same kinds and results as
 * fr_reduce, bit for bit; the block partial sums go to a fixed device array
 * of FR_NOALLOC_SCRATCH_DOUBLES doubles (512 KiB, one per GPU, held until the
 * process ends) instead of the context's buffer, so meshes of at most
 * (FR_NOALLOC_SCRATCH_DOUBLES - 8) / 5 * 256 elements (3,354,624) are
 * accepted and larger ones get FR_EINVAL.  Calls that overlap on one GPU
 * share that array and race. */
#define FR_NOALLOC_SCRATCH_DOUBLES 65536
int fr_reduce_noalloc(fr_ctx *ctx, int kind, double *out);

/* Error norms of U_spts: solver::integrateError (solver.cpp:1507-1574) with
 * calcError (funcs.cpp:1875-2056), the numbers writeError prints
 * (output.cpp:1055-1123, testCase 1).  The fields mirror the `input` members
 * they read (input.cpp:224-248, 276, 379-380). */
typedef struct fr_error_params {
  int testCase;   /* 0: the solution itself is integrated; 1: its error against the analytic solution */
  int errorNorm;  /* 0: |integral of the error - integral of U| at the solution points; 1 L1; 2 L2 */
  int quadOrder;  /* errorNorm 1, 2: (quadOrder+1)^nDims Gauss points per element, 0..10 (reference default 10) */
  int icType;     /* Euler / NS: 0 free-stream entropy, 1 isentropic vortex, 2 Liang's vortex;
                     advection: 0 Gaussian, 1 sine, 2 cosine product */
  double time;    /* the flow time of the analytic solution */
  double rhoIC, pIC;                /* icType 0 (Euler / NS): the free stream */
  double vortexAngle;               /* icType 2 (Euler / NS) */
  double xmin, xmax, ymin, ymax;    /* the periodic box of the analytic solutions (calcError's
                                       CREATE_MESH box, or its 'standard' box) */
} fr_error_params;
/* out[nFields]; quads and hexes (FR_EINVAL for triangles) */
int fr_integrate_error(fr_ctx *ctx, const fr_error_params *p, double *out);

/* Time step.  fr_set_dt: fixed params->dt (dtType 0).  fr_calc_dt: CFL time
 * step from the current U_spts -- ele::calcWaveSpFpts (ele.cpp:1031-1084,
 * static grids), ele::calcDt (1086-1105), solver::calcDt (solver.cpp:293-305);
 * dtType 1: params->dt = min over elements; dtType 2: every element advances
 * with its own dt (timeStepA/B with eles[e]->dt, solver.cpp:307-385) and
 * params->dt = the minimum.  As the reference does (calcWaveSpFpts at the
 * first iteration only, solver.cpp:293-297, 1497-1503; the Rusanov face flux
 * storing the wave speeds, face.cpp:354; calcDt at the start of every update,
 * solver.cpp:221-222), the first fr_update after this call takes the wave
 * speeds from the state and every later one from those its stage kernels'
 * face fluxes stored, so dt follows the solution.  Triangles (no reference
 * counterpart) use the same formula with their dense traces, the element area
 * as vol and the quads' CFL limit.  fr_get_dt: the dt each element advances
 * with, in mesh element order. */
int fr_set_dt(fr_ctx *ctx, double dt);
int fr_calc_dt(fr_ctx *ctx, int dtType, double CFL, double *dtMin);
int fr_get_dt(fr_ctx *ctx, double *dtEle /* nEles */);

/* Adjoint (no reference counterpart: FlurryPP has no adjoint).  FP64
 * contexts (FR_EINVAL otherwise): Euler, Navier-Stokes (LDG), advection and
 * advection-diffusion.  Arrays in the reference layout.
 * fr_vjp_residual: out = (d divF / dU)^T lam at the current U_spts
 *   (the vector-Jacobian product of solver::calcResidual).
 * fr_vjp_step: lam0 = (d U_{n+1} / d U_n)^T lam1 for one solver::update
 *   from the current U_spts (stage states recomputed; U_spts unchanged).
 * fr_vjp_steps: nSteps updates from the current U_spts (which ends at U_N,
 *   as fr_update(nSteps)), then the reverse sweep lam0 = (dU_N/dU_0)^T lamN,
 *   keeping a state every `every` steps and recomputing each segment.
 * fr_reduce_gradient(kind 10): d(kinetic energy)/dU_spts, for objectives
 *   J = KE(U_N). */
int fr_vjp_residual(fr_ctx *ctx, const double *lam, double *out);
int fr_vjp_step(fr_ctx *ctx, const double *lam1, double *lam0);
int fr_vjp_steps(fr_ctx *ctx, int nSteps, int every, const double *lamN, double *lam0);
int fr_reduce_gradient(fr_ctx *ctx, int kind, double *out);

/* Batches (no reference counterpart): many contexts of one equation,
 * precision and timeType (sizes, orders, boundaries, and quads, triangles and
 * hexes mixed; inviscid) advanced together, one persistent launch per RK
 * stage, element type and order over all their element blocks, one step's
 * launches replayed from a captured hipGraph.
 * fr_batch_create autotunes, per group (element type and order), the tile
 * shape (threads per element and elements per workgroup) and the persistent
 * grid on residual evaluations timed group by group (which overwrite
 * divF_spts[0] only).  fr_batch_tuning reports the largest grid, the work
 * items (element blocks) and the sum of the groups' stage times;
 * fr_batch_group_tuning one group's choice (group = P for quads, 5 + P for
 * triangles, 10 + P for hexes; nItems 0 when the batch has no such mesh).
 * fr_batch_update is fr_update(nSteps) on every member; the contexts stay
 * usable on their own between batch calls. */
typedef struct fr_batch fr_batch;
int fr_batch_create(fr_ctx *const *ctxs, int n, fr_batch **out);
int fr_batch_update(fr_batch *batch, int nSteps);
int fr_batch_tuning(fr_batch *batch, int *grid, int *nItems, double *msStage0);
int fr_batch_group_tuning(fr_batch *batch, int group, int *threadsPerElement, int *elementsPerBlock, int *grid,
                          int *nItems, double *ms);
/* For the batches created after it (process-wide): -1 lets the tuner choose
 * the tile layout (the default), 0 fixes the per-mesh kernels' tile (a power
 * of two threads per element), 1 the packed tile of 2D groups (as many
 * threads per element as it has solution or flux points, whichever is more;
 * groups where both coincide keep layout 0).  The grid is tuned either way. */
int fr_batch_force_tile(int layout);
int fr_batch_destroy(fr_batch *batch);

/* Stream control.  fr_set_stream(ctx, (void*)hipStream_t) makes the context
 * launch on the caller's stream (NULL = the context's own stream). */
int fr_set_stream(fr_ctx *ctx, void *stream);
int fr_synchronize(fr_ctx *ctx);

/* Solution files (writeData, output.cpp:11-24).  plotType 1: ParaView
 * <dataFileName>_<iter %09d>.vtu (writeParaview, output.cpp:103-391; 3D
 * only, FR_EINVAL otherwise, as there): per element the density, velocity
 * and pressure at the (P+3)^3 plot points (the tensor grid of -1, the
 * solution points and +1; extrapolateUPpts, solver.cpp:449-484), their
 * positions and the (P+2)^3 linear sub-hexes, with the TIME and ITER
 * comments.  plotType 0: <dataFileName>.csv.<iter %09d> (writeCSV,
 * output.cpp:26-100): position and primitives at every solution point,
 * then every flux point.  Elements in mesh order. */
int fr_write_data(fr_ctx *ctx, const char *dataFileName, int plotType, int iter, double time);

/* Restart from <dataFileName>_<restartIter %09d>.vtu (solver::readRestartFile,
 * solver.cpp:1420-1482; ele::restart, ele.cpp:1451-1880): density, velocity
 * and pressure at the solution points -> U_spts, interpolated when the
 * file's order differs from this context's.  *time (may be NULL) receives
 * the file's TIME comment and is left unchanged when there is none.
 * Errors: FR_EIO for a missing file, no <UnstructuredGrid> tag, or a
 * missing / short element (the reference fills such an element with 100). */
int fr_read_restart(fr_ctx *ctx, const char *dataFileName, int restartIter, double *time);

/* Exact binary checkpoint of U_spts (FP64, reference layout) with the
 * iteration and time; reading checks nDims, order, nFields and nEles.  No
 * reference counterpart (its restart goes through the 16-digit .vtu). */
int fr_write_checkpoint(fr_ctx *ctx, const char *path, int iter, double time);
int fr_read_checkpoint(fr_ctx *ctx, const char *path, int *iter, double *time);

/* Geometry layout of the stage kernels.  request:
 *   -1  query;
 *    0  per-point arrays (metrics and detJ at every solution point, normal
 *       and dA at every flux point);
 *    1  compact (the default): on a mesh whose elements are all affine
 *       (metrics and detJ constant over each element, normal and dA constant
 *       over each face, to 1e-13 relative) one metric set per element and one
 *       normal per face, else per point;
 *    2  nodal: on inviscid meshes of linear quads / hexes (4 / 8 nodes, any
 *       shape) the stage kernels form the metrics and detJ at each solution
 *       point and the normal and dA at each flux point from the element's
 *       corner nodes (21 coefficients per hex) instead of reading them,
 *       else per point;
 *    3  as 2 for the solution-point metrics, with the stored face geometry.
 * *active (may be NULL) receives the layout in use (0..3).  No reference
 * counterpart: FlurryPP always stores metrics per point
 * (ele::calcTransforms, ele.cpp:202-342); the results agree to rounding. */
int fr_geometry_mode(fr_ctx *ctx, int request, int *active);

/* keep = 1 (the default): the final stage of fr_update stores its
 * divergence in divF_spts[nRK-1], as solver::update leaves it
 * (solver.cpp:212-250).  keep = 0: a multi-stage update does not store it --
 * nothing in the update reads it, and leaving it out saves one state-sized
 * write per step; fr_get_residual(ctx, nRK-1, ...) then fails with FR_EINVAL
 * until a residual of that step is computed again (fr_calc_residual).
 * divF_spts[0], which the residual norms read, is always stored.  No
 * reference counterpart. */
int fr_keep_final_residual(fr_ctx *ctx, int keep);

/* Kernel timing: with timing on, every launch of the residual kernel is
 * bracketed by HIP events on the launch stream; fr_kernel_times returns the
 * accumulated milliseconds and launch count since the last reset for
 * kernel id 0 = residual, 1 = trace, 2 = RK update, 3 = gradient (viscous). */
int fr_set_timing(fr_ctx *ctx, int on);
int fr_kernel_times(fr_ctx *ctx, int kernel, double *ms, long *count);

/* Algorithmic HBM bytes of one residual evaluation (the B_res model: state in
 * and residual out, per-spt adjugate metrics, one-sided normal + dA per face
 * flux point, 20-byte face record), in the context's precision. */
int fr_residual_bytes(fr_ctx *ctx, double *bytes);

#ifdef __cplusplus
}
#endif
#endif
