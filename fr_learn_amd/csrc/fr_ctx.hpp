// fr_ctx.hpp -- the library's internal context (one per fr_create) and batch,
// shared by the translation units of libfrhip: the ABI and host logic
// (fr_api.hip), the stage-kernel launches (fr_stage64.hip / fr_stage32.hip),
// the adjoint (fr_adjoint.hip) and the batched launches (fr_batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/fr.h"
#include "fr_batch.hpp"
#include "fr_host.hpp"
#include "fr_kernels.hpp"

namespace frimpl {

inline thread_local std::string g_err;

inline int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) return fail(FR_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename F>
inline bool dispatch(int D, int P, int NF, F &&f) {
#define FR_CASE(d, p, nf)                                                                          \
  if (D == d && P == p && NF == nf) {                                                              \
    f(std::integral_constant<int, d>{}, std::integral_constant<int, p>{}, std::integral_constant<int, nf>{}); \
    return true;                                                                                   \
  }
  FR_CASE(3, 1, 5) FR_CASE(3, 2, 5) FR_CASE(3, 3, 5) FR_CASE(3, 4, 5) FR_CASE(3, 5, 5)
  FR_CASE(3, 1, 1) FR_CASE(3, 2, 1) FR_CASE(3, 3, 1) FR_CASE(3, 4, 1) FR_CASE(3, 5, 1)
  FR_CASE(2, 1, 4) FR_CASE(2, 2, 4) FR_CASE(2, 3, 4) FR_CASE(2, 4, 4) FR_CASE(2, 5, 4)
  FR_CASE(2, 1, 1) FR_CASE(2, 2, 1) FR_CASE(2, 3, 1) FR_CASE(2, 4, 1) FR_CASE(2, 5, 1)
#undef FR_CASE
  return false;
}

inline unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace frimpl

struct fr_batch;

struct fr_ctx {
  fr_config cfg{};
  int prec = FR_PREC_FP64;
  int dev = 0;  // the HIP device the context was created on
  frhost::Operators ops;
  frhost::Connectivity conn;
  int D = 3, P = 3, NS = 0, NFP = 0, NFPF = 0, NFACE = 0, NF = 0, nE = 0, nRK = 0;
  double RKa[5] = {0}, RKb[4] = {0};
  frdev::Phys ph{};
  hipStream_t own = nullptr, stream = nullptr;
  // device buffers
  double *U = nullptr, *U0 = nullptr, *detJ = nullptr, *JG64 = nullptr, *fgeo64 = nullptr, *ownN = nullptr;
  double *wts = nullptr, *oned64 = nullptr, *part = nullptr;
  float *U32 = nullptr, *JG32 = nullptr, *fgeo32 = nullptr, *oned32 = nullptr;
  // compact geometry of affine meshes (frdev::GeoAddr): per element / face
  double *JGc64 = nullptr, *detJc = nullptr, *fgeoc64 = nullptr;
  float *JGc32 = nullptr, *fgeoc32 = nullptr;
  bool affineMesh = false;  // every element affine (compact arrays built)
  // nodal geometry of linear quads / hexes (frdev::LinearMap): the map
  // coefficients per element, from which the stage kernel forms its metrics
  double *geoN64 = nullptr;
  float *geoN32 = nullptr;
  // requested stage-kernel geometry (fr_geometry_mode): 0 per point,
  // 1 compact when affine, 2 nodal, 3 nodal metrics with stored face geometry
  int geoReq = 1;
  // error norms (fr_integrate_error): point tables and the element nodes
  double *errTab = nullptr, *Xd = nullptr;
  size_t errTabCap = 0;
  void *Us = nullptr;                 // RK stage state (residual precision)
  void *Uf[2] = {nullptr, nullptr};   // face traces, double-buffered
  int cur = 0;                        // Uf[cur] holds the traces of `traceOf`
  int traceOf = -1;                   // 0: U_spts, 1: stage state Us, -1: stale
  // element-block size of the launches that wrote Uf[cur] (their in-block
  // faces are missing from it), 0: every face written
  int traceEpb = 0;
  int revNext = 0;  // block order of the next stage launch (FR_STAGE_REVERSE)
  int keepFinal = 1;                  // the final RK stage stores divF_spts[nRK-1] (fr_keep_final_residual)
  bool finalDropped = false;          // divF[nRK-1] is stale: the last update did not store it
  void *divF[4] = {nullptr, nullptr, nullptr, nullptr};
  void *dUf = nullptr;                // viscous: gradient traces [ele][dim][field][fpt]
  void *Fv = nullptr;                 // viscous: common viscous normal flux [ele][field][fpt]
  double *dtG = nullptr;              // the global time step (one value)
  double *dtE = nullptr;              // per-element time steps (local time stepping)
  int dtType = 0;                     // 0 fixed dt, 1 global CFL dt, 2 local CFL dt
  // CFL time stepping (dtType != 0; solver::calcDt, solver.cpp:293-305): the
  // wave speeds waveSp[ele][fpt] (the reference's waveSp_fpts), the per-element
  // numerators CFL * limit(P) * 2 * vol, the surface weights, block minima
  double *waveSp = nullptr, *dtNum = nullptr, *wsurf = nullptr, *dtPart = nullptr;
  double cfl = 0;         // the CFL number of fr_calc_dt
  bool dtFirst = false;   // the next update is the first since fr_calc_dt (wave speeds from the state)
  bool dtReady = false;   // dtG / dtE already hold the next update's time step
  bool dtFrozen = false;  // the adjoint's replay: dt given, wave speeds untouched
  unsigned gen = 0;       // bumped by every setting the batched launch arguments capture
  std::vector<fr_batch *> batches;  // batches this context is a member of
  // adjoint work buffers (FP64, allocated on first use)
  double *adjA = nullptr, *adjTr = nullptr, *adjTr2 = nullptr;  // face cotangents, trace scratch
  double *adjLam = nullptr, *adjOut = nullptr, *adjNu = nullptr, *adjW[2] = {nullptr, nullptr};
  double *adjQf = nullptr, *adjB = nullptr;                     // viscous: gradient traces, jump cotangents
  double *adjUo = nullptr, *adjQo = nullptr, *adjCj = nullptr;  // viscous: face pull-backs, forward jumps
  double *adjS[4] = {nullptr, nullptr, nullptr, nullptr};       // RK stage states
  int32_t *fnbr = nullptr;
  int8_t *finfo = nullptr;
  size_t partCap = 0;
  std::vector<double> X;  // host copy of element node coordinates [device ele][node][3]
  std::vector<int> perm;  // device element -> mesh element (blocks of EPB face-connected elements)
  std::vector<int> iperm; // mesh element -> device element
  int epb = 1;
  bool tri = false;  // linear triangles (fr_tri_ops.cpp / fr_tri.hpp)
  // timing
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[4];
  double ms[4] = {0, 0, 0, 0};
  long cnt[4] = {0, 0, 0, 0};

  size_t nState() const { return (size_t)nE * NF * NS; }
  // the geometry layout the stage kernels use (the values of geoReq): the
  // request where the context supports it, else the per-point arrays
  int geoActive() const {
    if (geoReq == 1 && affineMesh) return 1;
    if ((geoReq == 2 || geoReq == 3) && geoN64 && !cfg.viscous && !tri) return geoReq;
    return 0;
  }
  size_t tsize() const { return prec == FR_PREC_FP32 ? 4 : 8; }
};

// Many contexts advanced together: one persistent launch per RK stage over
// all their element blocks, the launches of one step captured in a hipGraph.
struct fr_batch {
  std::vector<fr_ctx *> ctx;
  std::vector<unsigned> gen;  // each member's settings generation the arguments were built at
  std::vector<bool> alive;    // members not yet destroyed
  bool broken = false;        // a member was destroyed
  int nRK = 1, prec = FR_PREC_FP64;
  hipStream_t stream = nullptr;
  void *args[2] = {nullptr, nullptr};  // StageArgs per [trace parity][stage][instance]
  void *argsTune = nullptr;            // residual-only (mode 0) StageArgs per instance, for tuning
  // work items grouped by kernel: group g = P (quads of order P), 5 + P
  // (triangles) or 10 + P (hexes), g in 1..15 (frimpl::batchGroup); one list
  // per tile layout
  // (frdev::ElemGeo / TriGeo L), since an item is a block of that layout's EPB
  // elements
  frdev::BatchItem *items[2] = {nullptr, nullptr};
  int itemStart[2][16] = {}, itemCount[2][16] = {};
  int groupNF[16] = {};  // fields of the group's meshes
  // per group, chosen at fr_batch_create by timing: tile layout, persistent
  // grid and the residual-only launch time
  int lay[16] = {}, grid[16] = {};
  double groupMs[16] = {};
  int nItems = 0;  // items of the chosen layouts
  double tunedMs = 0;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  hipStream_t ps[16] = {nullptr};  // one per graph branch (capture)
  hipEvent_t fork = nullptr, join[16] = {nullptr};
};

namespace frimpl {

template <typename T>
inline int dmalloc(T **p, size_t n) {
  hipError_t e = hipMalloc((void **)p, (n ? n : 1) * sizeof(T));
  if (e != hipSuccess) return fail(FR_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return 0;
}

// Kernel timing bracket on the launch stream.
struct TimeScope {
  fr_ctx *c;
  int id;
  hipEvent_t a = nullptr, b = nullptr;
  TimeScope(fr_ctx *c_, int id_) : c(c_), id(id_) {
    if (c->timing) {
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      (void)hipEventRecord(a, c->stream);
    }
  }
  ~TimeScope() {
    if (c->timing) {
      (void)hipEventRecord(b, c->stream);
      c->ev[id].push_back({a, b});
    }
  }
};

// Time step source of the update kernels: dtE[ele * stride].
inline std::pair<const double *, int> dtSource(const fr_ctx *c) {
  if (c->dtType == 2) return {c->dtE, 1};
  return {c->dtG, 0};
}

// The mode and RK coefficient of stage s of one update (solver::update).
inline void stageMode(const fr_ctx *c, int s, int &mode, double &rk) {
  if (s + 1 < c->nRK) {
    mode = 1;
    rk = c->RKa[s + 1];
  } else if (c->cfg.timeType < 5) {
    mode = 2;
    rk = 0.0;
  } else {
    mode = 3;
    rk = c->RKa[c->nRK];
  }
}

// Launch of the stage kernel variant for `mode` (fr_stage64.hip / fr_stage32.hip).
template <typename T>
int launchStageArgs(fr_ctx *c, const frdev::StageArgs<T> &a, int mode);
// The residual evaluation of arguments `a` (mode 0, FP64, quads and hexes)
// with its debug outputs (a.dbgF, a.dbgFn, a.dbgJ; on the viscous path the
// gradient at the spts into dUs), fr_stage_dbg.hip.
int launchStageDbg(fr_ctx *c, const frdev::StageArgs<double> &a, double *dUs);
// (d divF / dU)^T lam at the state Us, device arrays (fr_adjoint.hip).
int vjpResidualDev(fr_ctx *c, const double *Us, const double *lam, double *out);
// triangle kernels (fr_tri.hip)
template <typename T>
int launchTriStage(fr_ctx *c, const frdev::StageArgs<T> &a, int mode);
template <typename T>
int launchTriTrace(fr_ctx *c, const T *U, T *Uf, const T *tab);
// CFL time steps of the triangles into c->dtE (k_tri_dt), on stream st
int launchTriDt(fr_ctx *c, const double *wsurf, double factor, hipStream_t st);
// c->part holds at least n doubles (the block partial sums of the
// reductions), allocated on first use (fr_api.hip).
int needPart(fr_ctx *c, size_t n);
// The time step of the next update into dtG / dtE on stream st (fr_api.hip):
// from the state on the first update after fr_calc_dt (and always for
// triangles), else from the wave speeds the last residual stored.
int dtPrepare(fr_ctx *c, hipStream_t st);
// Stage s of one batched step (fr_batch.hip); parity < 0 is the tuning launch,
// Ponly != 0 restricts it to the meshes of that order.
int batchLaunchGroup(fr_batch *b, int parity, int s, hipStream_t st, int grp, int lay, int grid);
int batchGroup(const fr_ctx *c);
// Element type of batch group g: nDims, order and whether triangles.
void batchGroupKind(int g, int &D, int &P, bool &tri);
int batchEpb(const fr_batch *b, const fr_ctx *c, int lay);

}  // namespace frimpl
