/*
 * fr_oracle.h -- CPU FP64 oracle for the flux-reconstruction residual + RK path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the FlurryPP
 * (fr_learn) algorithm, written to check the HIP product path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
 * as the checker (or as the timed CPU baseline), never as the measured or
 * shipped path.
 *
 * Layout follows the reference exactly (src/matrix.cpp:260-272, last index
 * fastest):  U_spts(spt, ele, field) = U[field + nFields*(ele + nEles*spt)],
 * F_spts(dim, spt, ele, field), JGinv_spts(dimRef, spt, ele, dimPhys), ...
 * Every contraction is a dense row-major dgemm over that layout, as the
 * reference does with cblas_dgemm (src/solver.cpp:401-412, 690-806).
 */
#ifndef FR_ORACLE_H
#define FR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* enums restated from include/global.hpp:41-81 */
enum { OR_TRI = 0, OR_QUAD = 1, OR_HEX = 4 };
enum { OR_ADVECTION_DIFFUSION = 0, OR_NAVIER_STOKES = 1 };
enum {
  OR_BC_NONE = -1, OR_PERIODIC = 0, OR_CHAR_INOUT = 1, OR_SUP_IN = 2,
  OR_SUP_OUT = 3, OR_SUB_IN = 4, OR_SUB_OUT = 5, OR_SLIP_WALL = 8,
  OR_ISOTHERMAL_NOSLIP = 9, OR_ADIABATIC_NOSLIP = 10, OR_SYMMETRY = 12,
  OR_ISOTHERMAL_NOSLIP_MOVING = 13
};
enum { OR_DG = 0, OR_SD = 1, OR_HU = 2, OR_CPLUS = 3 };

/* Subset of `input` (include/input.hpp) that the residual/RK path reads.
 * Field order is shared with the product's fr_config so one Python
 * description fills both. */
typedef struct or_params {
  int equation;    /* 0 advection-diffusion, 1 Euler/Navier-Stokes */
  int nDims;       /* 2 (quads) or 3 (hexes) */
  int order;       /* polynomial order P */
  int riemannType; /* 0 Rusanov, 1 Roe (2D only) */
  int viscous;
  int timeType;    /* 0 FE, 2 RK2, 4 RK44 ("RK45"), 5, 6 Jameson-style */
  int sptsType;    /* 0 Gauss-Legendre, 1 Gauss-Lobatto */
  int vcjhScheme;  /* 0 DG, 1 SD, 2 HU, 3 C+ */
  int fixVis;
  int viscFix;     /* 0: reproduce reference viscous-flux index defects; 1: corrected */
  double dt;
  double gamma, advectVx, advectVy, advectVz, lambda, diffD;
  double rhoBound, uBound, vBound, wBound, pBound;
  double penFact, tau, mu_inf, prandtl, rt_inf, c_sth;
  double TWall, RGas, uWall, vWall, wWall;
} or_params;

typedef struct or_solver or_solver;

/* Build operators, geometry and faces.
 *  xv[nVerts][3]              vertex coordinates
 *  c2v[nEles][nNodes]         element -> vertex (geometry), Gmsh corner order
 *  vmap[nVerts] or NULL       canonical vertex id used for face matching
 *                             (periodic identification); identity if NULL
 *  bcType[nBounds], bndOff[nBounds+1], bndPts[...]: vertex sets per boundary
 * Returns NULL on error (or_last_error()). */
or_solver *or_create(const or_params *p, int nVerts, const double *xv,
                     int nEles, int nNodes, const int *c2v, const int *vmap,
                     int nBounds, const int *bcType, const int *bndOff,
                     const int *bndPts);
void or_destroy(or_solver *s);
const char *or_last_error(void);

void or_sizes(const or_solver *s, int *out8); /* nSpts nFpts nEles nFields nDims nFaces nIntFaces nRK */

/* dense operators: which = 0 spts_to_fpts (nFpts x nSpts), 1 grad_spts
 * (nDims x nSpts x nSpts), 2 extrapolateFn (nDims x nFpts x nSpts),
 * 3 correction (nSpts x nFpts), 4 correctU (nDims x nSpts x nFpts),
 * 5 loc_spts (nSpts x 3), 6 loc_fpts (nFpts x 3), 7 spts1D (P+1),
 * 8 quadrature weights (nSpts) */
int or_get_operator(const or_solver *s, int which, double *out);

/* geometry: 0 JGinv_spts, 1 detJac_spts, 2 norm_fpts, 3 dA_fpts, 4 pos_spts,
 * 5 Jac_spts (all reference layout) */
int or_get_geometry(const or_solver *s, int which, double *out);

/* faces: per face (L ele, L local face, R ele, R local face, relRot, bcType)
 * in out6[nFaces*6]; fptR[nFaces*nFptsFace] right-element fpt per left fpt */
int or_get_faces(const or_solver *s, int *out6, int *fptR);

void or_set_state(or_solver *s, const double *U);
void or_get_state(const or_solver *s, double *U);
/* arrays: 0 U_spts, 1 U_fpts, 2 F_spts, 3 Fn_fpts, 4 disFn_fpts,
 * 5 dU_spts, 6 dU_fpts, 7 dUc_fpts, 8 U0, 9 waveSp_fpts [fpt][ele] (after
 * or_calc_dt), 10+s divF_spts[s] */
int or_get_array(const or_solver *s, int which, double *out);

void or_calc_residual(or_solver *s, int step);
void or_update(or_solver *s);
void or_time_step_a(or_solver *s, int step, double rkval);
void or_time_step_b(or_solver *s, int step);
void or_copy_u_u0(or_solver *s);
void or_copy_u0_u(or_solver *s);
/* residual norm of divF_spts[0] (ele.cpp:1882-1903 summed as output.cpp:846-892) */
void or_norm_residual(const or_solver *s, int normType, double *out);
void or_set_threads(int n);
/* Error norms (solver::integrateError, solver.cpp:1507-1574; calcError,
 * funcs.cpp:1875-2056).  out[nFields]. */
typedef struct or_error_params {
  int testCase, errorNorm, quadOrder, icType;
  double time, rhoIC, pIC, vortexAngle, xmin, xmax, ymin, ymax;
} or_error_params;
int or_integrate_error(const or_solver *s, const or_error_params *ep, double *out);
/* test hooks: shape functions of an nNodes element at r[3] (sh[nNodes],
 * dsh[nNodes][3], either may be NULL), and the lattice indices of the nodes
 * of the nSide^3 Lagrange hex in node order (abc[n][3]; returns the count) */
void or_shape(int nDims, int nNodes, const double *r, double *sh, double *dsh);
int or_hex_lattice(int nSide, int *abc);

/* pointwise test hooks (flux.cpp / face.cpp / boundFace.cpp / polynomials.cpp
 * restatements), so tests can pin them against textbook formulas */
double or_calc_dt(or_solver *s, int dtType, double CFL);
double or_get_dt(const or_solver *s, int e);
void or_point_inviscid_flux(const or_params *p, const double *U, double *F /* [3][5] */);
void or_point_common_flux(const or_params *p, int kind, const double *UL, const double *UR, const double *n,
                          double *Fn); /* kind 0 Rusanov, 1 Roe, 2 Lax-Friedrichs */
void or_point_boundary_state(const or_params *p, int bc, const double *UL, const double *n, double *UR);
void or_point_viscous_flux(const or_params *p, const double *U, const double *grad, double *F);
void or_point_viscous_common(const or_params *p, int isBnd, const double *UL, const double *UR, const double *gL,
                             const double *gR, const double *n, double *Fn);
void or_point_viscous_bc_gradient(const or_params *p, int bc, const double *UL, const double *UR, const double *gL,
                                  double *gR);
double or_point_dvcjh(double r, int mode, int order, double eta);
double or_point_eta(int scheme, int order);

#ifdef __cplusplus
}
#endif
#endif
