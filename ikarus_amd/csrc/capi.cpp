// capi.cpp — the extern "C" boundary plus the host-side NewtonRaphson / LoadControl loops.
//
// The loops restate NewtonRaphson::solve (ikarus/solver/nonlinearsolver/newtonraphson.hh:196-257) and
// LoadControl::run (ikarus/controlroutines/loadcontrol.inl:21-57) with the iterate, residual and
// tangent resident in device memory: per Newton iteration the host reads back two scalars (||R||
// and the PCG state), nothing else.
#include <algorithm>
#include <cmath>
#include <exception>

#include "ik_internal.hh"

using ik::Error;
using ik::require;

namespace {

template <class F>
int guard(ik_handle* h, F&& f) {
  try {
    f();
    if (h) h->err.clear();
    return IK_OK;
  } catch (const Error& e) {
    if (h) h->err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    if (h) h->err = e.what();
    return IK_EINVAL;
  }
}

__global__ void negateKernel(double* __restrict__ x, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) x[i] = -x[i];
}
__global__ void gatherKernel(const double* __restrict__ src, const int32_t* __restrict__ idx, int64_t n,
                             double* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) dst[i] = src[idx[i]];
}

// duN1 <- duN ; duN <- u - uPrev ; uPrev <- u
__global__ void historyKernel(const double* __restrict__ u, double* __restrict__ uPrev, double* __restrict__ duN,
                              double* __restrict__ duN1, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double cur = u[i];
    duN1[i]          = duN[i];
    duN[i]           = cur - uPrev[i];
    uPrev[i]         = cur;
  }
}

// u[d] = lambda ubar[d] on the constrained DOFs where that value is nonzero: the update with the inhomogeneous
// boundary values (nonlinearsolverfactory.hh:45-54, dirichletvalues.hh:214-257)
__global__ void syncIdbcKernel(const int32_t* __restrict__ fl, int64_t nf, const double* __restrict__ ubar,
                               double lambda, double* __restrict__ u) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nf; i += (int64_t)gridDim.x * 256) {
    const int32_t d = fl[i];
    const double v  = ubar[d] * lambda;
    if (v != 0.0) u[d] = v;
  }
}
__global__ void scaleKernel(double* __restrict__ x, double a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) x[i] *= a;
}
__global__ void zeroListKernel(const int32_t* __restrict__ fl, int64_t nf, double* __restrict__ f) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nf; i += (int64_t)gridDim.x * 256) f[fl[i]] = 0.0;
}

inline int grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  return (int)std::min<int64_t>(std::max<int64_t>(g, 1), 4096);
}

// ---- plane strain as a one-layer extrusion (see ik_handle::plane2d) --------------------------------------
constexpr double kSqrt2 = 1.4142135623730950488;

// b (3D, both layers) <- s a (2D, FlatInterleaved 2v + k); the z components are zero
__global__ void expand2dKernel(const double* __restrict__ a, int64_t nn2, double s, double* __restrict__ b) {
  for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < nn2; v += (int64_t)gridDim.x * 256) {
    const double x = a[2 * v] * s, y = a[2 * v + 1] * s;
    double* b0 = b + 3 * v;
    double* b1 = b + 3 * (v + nn2);
    b0[0] = x;
    b0[1] = y;
    b0[2] = 0.0;
    b1[0] = x;
    b1[1] = y;
    b1[2] = 0.0;
  }
}
// a (2D) <- s0 (layer 0 of b) + s1 (layer 1 of b)
__global__ void contract2dKernel(const double* __restrict__ b, int64_t nn2, double s0, double s1,
                                 double* __restrict__ a) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < 2 * nn2; i += (int64_t)gridDim.x * 256) {
    const int64_t v = i >> 1, k = i & 1;
    a[i] = s0 * b[3 * v + k] + s1 * b[3 * (v + nn2) + k];
  }
}

// 3D device vector dst <- s * (2D vector src, host or device)
void expandIn(ik_handle* h, const double* src, int where, double s, double* dst) {
  const int64_t n2 = 2 * h->nn2;
  if (where != IK_DEVICE) {
    h->map2d.ensure(n2);
    IK_HIP(hipMemcpyAsync(h->map2d.p, src, n2 * sizeof(double), hipMemcpyHostToDevice, h->stream));
    src = h->map2d.p;
  }
  hipLaunchKernelGGL(expand2dKernel, dim3(grid(h->nn2)), dim3(256), 0, h->stream, src, h->nn2, s, dst);
}
// 2D vector dst (host or device) <- s0 layer 0 + s1 layer 1 of the 3D device vector src
void contractOut(ik_handle* h, const double* src, double s0, double s1, double* dst, int where) {
  const int64_t n2 = 2 * h->nn2;
  h->map2d.ensure(n2);
  hipLaunchKernelGGL(contract2dKernel, dim3(grid(n2)), dim3(256), 0, h->stream, src, h->nn2, s0, s1, h->map2d.p);
  IK_HIP(hipMemcpyAsync(dst, h->map2d.p, n2 * sizeof(double),
                        where == IK_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, h->stream));
  IK_HIP(hipStreamSynchronize(h->stream));
}

// owning storage of the extruded description
struct Extruded {
  std::vector<double> X;
  std::vector<int32_t> conn;
  std::vector<uint8_t> fixed;
  std::vector<int64_t> pdof;
  std::vector<double> pcoef;
};

// The 3D description of a plane-strain Q1 problem: layer l of 2D node v is node v + l nn (z = l sqrt(2)); hex
// corner a (Dune order) is quad corner a & 3 on layer a >> 2; u_z is fixed on every node; an edge load on the
// quad edge normal to axis 0 or 1 is the traction on the hex face normal to the same axis; a 2D point load puts
// coef / sqrt(2) on both layers.
ik_problem_desc extrude2d(ik_handle* h, const ik_problem_desc& d, Extruded& ex) {
  require(d.num_nodes > 0 && d.num_elements > 0 && d.X && d.conn, IK_EINVAL, "empty mesh");
  require(d.eas_modes == 0, IK_ENOTIMPL, "plane strain on the device has no EAS modes");
  require(d.num_nodes * 6 < (int64_t)INT32_MAX, IK_EINVAL, "too many DOFs for int32 indices");
  const int64_t nn = d.num_nodes, ne = d.num_elements;
  h->plane2d = true;
  h->nn2     = nn;
  ex.X.resize((size_t)nn * 6);
  for (int64_t v = 0; v < nn; ++v)
    for (int l = 0; l < 2; ++l) {
      double* x = &ex.X[(size_t)(v + l * nn) * 3];
      x[0]      = d.X[2 * v];
      x[1]      = d.X[2 * v + 1];
      x[2]      = l * kSqrt2;
    }
  h->hconn2.assign(d.conn, d.conn + ne * 4);
  ex.conn.resize((size_t)ne * 8);
  for (int64_t e = 0; e < ne; ++e)
    for (int a = 0; a < 8; ++a) {
      const int32_t v = d.conn[e * 4 + (a & 3)];
      require(v >= 0 && v < nn, IK_EINVAL, "connectivity index out of range");
      ex.conn[(size_t)e * 8 + a] = v + (int32_t)((a >> 2) * nn);
    }
  if (d.dirichlet)
    h->hfixed2.assign(d.dirichlet, d.dirichlet + nn * 2);
  else
    h->hfixed2.assign(nn * 2, 0);
  ex.fixed.resize((size_t)nn * 6);
  for (int64_t v = 0; v < nn; ++v)
    for (int l = 0; l < 2; ++l)
      for (int k = 0; k < 3; ++k) ex.fixed[(size_t)(v + l * nn) * 3 + k] = k < 2 ? h->hfixed2[2 * v + k] : 1;
  for (int64_t f = 0; f < d.num_faces; ++f)
    require(d.face_axis && d.face_axis[f] >= 0 && d.face_axis[f] < 2, IK_EINVAL, "bad traction edge of a 2D mesh");
  for (int64_t i = 0; i < d.num_point_loads; ++i) {
    require(d.point_dof && d.point_coef && d.point_dof[i] >= 0 && d.point_dof[i] < 2 * nn, IK_EINVAL,
            "point load dof out of range");
    const int64_t v = d.point_dof[i] / 2, k = d.point_dof[i] % 2;
    for (int l = 0; l < 2; ++l) {
      ex.pdof.push_back(3 * (v + l * nn) + k);
      ex.pcoef.push_back(d.point_coef[i] / kSqrt2);
    }
  }
  ik_problem_desc o = d;
  o.dim             = 3;
  o.num_nodes       = 2 * nn;
  o.X               = ex.X.data();
  o.conn            = ex.conn.data();
  o.dirichlet       = ex.fixed.data();
  o.traction[2]     = 0.0;
  o.num_point_loads = (int64_t)ex.pdof.size();
  o.point_dof       = ex.pdof.data();
  o.point_coef      = ex.pcoef.data();
  return o;
}

void copyIn(ik_handle* h, double* dst, const double* src, int64_t n, int where) {
  IK_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), where == IK_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                        h->stream));
}
void copyOut(ik_handle* h, double* dst, const double* src, int64_t n, int where) {
  IK_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), where == IK_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                        h->stream));
  IK_HIP(hipStreamSynchronize(h->stream));
}

void ensurePattern(ik_handle* h) {
  if (!h->patternBuilt) ik::buildPattern(h);
}

void requireBound(ik_handle* h) {
  require(h->bound, IK_ENOTBOUND, "The assembler is not bound to a requirement (call ik_set_requirement first)");
}

double normR(ik_handle* h) { return std::sqrt(ik::dot(h, h->R.p, h->R.p, h->ndof)); }

// residual at the bound requirement into h->R (FULL)
void assembleR(ik_handle* h, int dbc) {
  ensurePattern(h);
  ik::assembleKR(h, dbc, IK_FP64, false, true);
}
void assembleK(ik_handle* h, int dbc, int precision) {
  ensurePattern(h);
  ik::assembleKR(h, dbc, precision, true, false);
  h->kPrecision = precision;
  h->kDbc       = dbc;
  h->kValid     = true;
}
bool kReady(const ik_handle* h, int dbc, int precision) {
  return h->kValid && h->kDbc == dbc && h->kPrecision == precision;
}
// The residual at a new state (the start of a Newton solve, or the state a correction just reached). The EAS
// element's residual-only pass still forms
// D, L and Rt and eliminates (three quarters of the time of its R+K pass), so for EAS handles the pass also assembles
// the condensed K (FULL, the given precision): the next iteration solves with it instead of assembling again, and
// R is the same, bit for bit, as the residual-only pass gives. The plain element's residual pass is a tenth of
// its tangent, so there R is assembled alone.
void assembleRAfterUpdate(ik_handle* h, int precision, bool withK) {
  if (!(withK && h->neas)) {
    assembleR(h, IK_DBC_FULL);
    return;
  }
  ensurePattern(h);
  ik::assembleKR(h, IK_DBC_FULL, precision, true, true);
  h->kPrecision = precision;
  h->kDbc       = IK_DBC_FULL;
  h->kValid     = true;
}

void syncIdbc(ik_handle* h) {
  if (!h->hasUbar || h->nfixed == 0) return;
  hipLaunchKernelGGL(syncIdbcKernel, dim3(grid(h->nfixed)), dim3(256), 0, h->stream, h->fixedList.p, h->nfixed,
                     h->ubar.p, h->lambda, h->u.p);
}

// rhsNorm: ||rhs|| when the caller has it, negative otherwise
void linearSolve(ik_handle* h, const double* rhs, double* x, const ik_solver_opts* o, double rtol,
                 ik_solver_info* info, double rhsNorm = -1.0) {
  require(h->kValid && h->kDbc == IK_DBC_FULL, IK_EINVAL, "ik_solve needs a FULL-DBC matrix assembled first");
  if (o->kind == IK_PCG_JACOBI_FP64) {
    require(h->kPrecision == IK_FP64, IK_EINVAL, "FP64 PCG needs an FP64 tangent");
    int it     = 0;
    double rel = 0;
    ik::pcgJacobi64(h, rhs, x, rtol, o->max_iter, &it, &rel);
    info->iterations   = it;
    info->refinements  = 0;
    info->rel_residual = rel;
    info->converged    = rel <= rtol ? 1 : 0;
  } else if (o->kind == IK_PCG_BJACOBI_FP32_IR || o->kind == IK_PCG_TWOLEVEL_FP32_IR) {
    require(h->kPrecision == IK_FP32, IK_EINVAL, "mixed PCG needs an FP32 tangent");
    ik_solver_opts oo = *o;
    oo.rtol           = rtol;
    ik::pcgMixed(h, rhs, x, &oo, info, rhsNorm);
  } else {
    throw Error{IK_ENOTIMPL, "unknown linear solver kind"};
  }
}

void newtonSolve(ik_handle* h, const ik_nr_settings* s, double stepSize, ik_nr_info* info) {
  requireBound(h);
  ensurePattern(h);
  // the loop's assemblies leave their error-flag reads to the synchronisations the loop makes anyway (the PCG's
  // done-flag reads, the residual norm), so an invalid element is still reported from within this call
  struct DeferScope {
    ik_handle* h;
    ~DeferScope() { h->deferErr = false; }
  } deferScope{h};
  h->deferErr = true;
  const int64_t n = h->ndof;
  h->du.ensure(n);
  *info              = ik_nr_info{};
  info->success      = 1;
  const bool idbcF   = h->hasUbar && h->nfixed > 0 && stepSize != 0.0;
  if (idbcF) {
    // force of the inhomogeneous boundary values, K_raw(x) ubar with the constrained entries zeroed, added to
    // the first residual scaled by the step (newtonraphson.hh:215-218; functionhelper.hh:170-184)
    assembleK(h, IK_DBC_RAW, IK_FP64);
    h->idbcF.ensure(n);
    ik::spmv64(h, h->ubar.p, h->idbcF.p);
    hipLaunchKernelGGL(zeroListKernel, dim3(grid(h->nfixed)), dim3(256), 0, h->stream, h->fixedList.p, h->nfixed,
                       h->idbcF.p);
  }
  // the first residual; an EAS handle takes it from the R+K pass (its residual-only pass forms D, L and Rt and
  // eliminates as well, and R is the same bit for bit), and the first iteration solves with that K
  if (h->neas)
    assembleRAfterUpdate(h, s->precision, true);
  else
    assembleR(h, IK_DBC_FULL);
  if (idbcF) ik::axpy(h, stepSize, h->idbcF.p, h->R.p, n);
  double rNorm        = normR(h);
  if (!std::isfinite(rNorm)) {
    info->residual_norm = rNorm;
    info->success       = 0;
    return;
  }
  double rPrev        = rNorm;
  double etaPrev      = 0.0;
  info->residual_norm = rNorm;
  int iter            = 0;
  if (!h->neas) h->kValid = false;
  while ((rNorm > s->tol && iter < s->max_iter) || iter < s->min_iter) {
    if (!kReady(h, IK_DBC_FULL, s->precision)) assembleK(h, IK_DBC_FULL, s->precision);
    // K x = R ; correction = -x (solve(correction, -rx), newtonraphson.hh:222-223)
    double rtol = s->linear.rtol;
    if (s->forcing > 0.0) {
      // inexact Newton, Eisenstat-Walker choice 2 (gamma = forcing, eta <= eta_max) with the safeguard
      // eta >= gamma eta_prev^2 and a floor: no need to solve below half the absolute Newton tolerance
      const double etaMax = s->eta_max > 0.0 ? s->eta_max : 0.1;
      double eta = etaMax;
      if (iter > 0) {
        eta = s->forcing * (rNorm / rPrev) * (rNorm / rPrev);
        const double sg = s->forcing * etaPrev * etaPrev;
        if (sg > 0.1) eta = std::max(eta, sg);
      }
      eta     = std::min(eta, etaMax);
      eta     = std::max(eta, 0.5 * s->tol / std::max(rNorm, 1e-300));
      rtol    = std::max(eta, s->linear.rtol);
      etaPrev = rtol;
    }
    rPrev = rNorm;
    ik_solver_info li{};
    linearSolve(h, h->R.p, h->du.p, &s->linear, rtol, &li, rNorm);  // rNorm = ||R|| of this very R
    info->linear_iterations += li.iterations;
    hipLaunchKernelGGL(negateKernel, dim3(grid(n)), dim3(256), 0, h->stream, h->du.p, n);
    if (h->neas) ik::easUpdate(h, h->du.p);  // CORRECTION_UPDATED (newtonraphson.hh:230)
    ik::axpy(h, 1.0, h->du.p, h->u.p, n);
    syncIdbc(h);  // syncParameterAndGlobalSolution (newtonraphson.hh:236-237)
    h->kValid        = false;
    h->easCacheValid = false;
    assembleRAfterUpdate(h, s->precision, true);
    rNorm               = normR(h);
    info->residual_norm = rNorm;
    ++iter;
    if (!std::isfinite(rNorm)) break;
  }
  // the last correction's norm (reported only, so read once: du still holds it)
  if (iter > 0) info->correction_norm = std::sqrt(ik::dot(h, h->du.p, h->du.p, n));
  // success = iter != maxIter (newtonraphson.hh:251-252); a residual norm that is not finite is a failure here
  // (the reference's loop ends on it, since NaN > tol is false, and reports success)
  if (iter == s->max_iter || !std::isfinite(rNorm)) info->success = 0;
  info->iterations = iter;
  ik::finishErrCheck(h, true);
}

// ---- arc-length path following -------------------------------------------------------------------
struct ArcState {
  double Dl = 0.0, psi = 0.0, stepSize = 0.0;
  double ub2 = 0.0;        // |ubar|^2 of the inhomogeneous Dirichlet values (0 without them)
  bool predicted = false;  // the initial prediction has run: the subsidiary derivatives are live
};

// ArcLength::operator() (pathfollowingfunctions.hh:103-115) with the inhomogeneous Dirichlet increment of the step,
// idbcDelta = Dl ubar (idbcIncrement, controlroutines/common.hh:63-76): f = sqrt(DD.DD + Dl^2 |ubar|^2 + psi^2 Dl^2)
// - s; returns f, 1/root (the factor of df/dDD = DD/root) and df/dDl = (|ubar|^2 + psi^2) Dl / root, both zero
// before the initial prediction
static void arcSubsidiary(ik_handle* h, const ArcState& a, double& f, double& invRoot, double& dfdDl) {
  const double dl2  = a.Dl * a.Dl;
  const double root = std::sqrt(ik::dot(h, h->arcDD.p, h->arcDD.p, h->ndof) + dl2 * a.ub2 + a.psi * a.psi * dl2);
  f                 = root - a.stepSize;
  invRoot           = a.predicted ? 1.0 / root : 0.0;
  dfdDl             = a.predicted ? (a.ub2 + a.psi * a.psi) * a.Dl / root : 0.0;
}

// force of the inhomogeneous Dirichlet values at the bound state into h->idbcF: K_raw(u) ubar with the
// constrained entries zeroed (utils::obtainForcesDueToIDBC, functionhelper.hh:170-184); FP64 whatever the
// tangent precision
static void idbcForce(ik_handle* h) {
  assembleK(h, IK_DBC_RAW, IK_FP64);
  h->idbcF.ensure(h->ndof);
  ik::spmv64(h, h->ubar.p, h->idbcF.p);
  hipLaunchKernelGGL(zeroListKernel, dim3(grid(h->nfixed)), dim3(256), 0, h->stream, h->fixedList.p, h->nfixed,
                     h->idbcF.p);
  h->kValid = false;
}

// NewtonRaphsonWithSubsidiaryFunction::solve (newtonraphsonwithscalarsubsidiaryfunction.hh:190-307) at the
// bound requirement
static void arcNewton(ik_handle* h, const ik_nr_settings* s, ArcState& a, ik_nr_info* info) {
  const int64_t n  = h->ndof;
  const bool idbc  = h->hasUbar && h->nfixed > 0;
  *info            = ik_nr_info{};
  info->success    = 1;
  // Fext0 = -R(0, 1) (:212-221): u = 0 and lambda = 1 with the rest of the state as it is (EAS alpha kept)
  {
    h->tmp.ensure(n);
    IK_HIP(hipMemcpyAsync(h->tmp.p, h->u.p, n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    const double lam0 = h->lambda;
    ik::launchZero(h->u.p, n, h->stream);
    h->lambda = 1.0;
    assembleR(h, IK_DBC_FULL);
    IK_HIP(hipMemcpyAsync(h->arcF.p, h->R.p, n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    hipLaunchKernelGGL(negateKernel, dim3(grid(n)), dim3(256), 0, h->stream, h->arcF.p, n);
    IK_HIP(hipMemcpyAsync(h->u.p, h->tmp.p, n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    h->lambda = lam0;
  }
  // an EAS handle takes the residual from the R+K pass, and the first iteration solves with its K (with
  // inhomogeneous Dirichlet values every iteration assembles K_raw first and then K)
  if (h->neas && !idbc)
    assembleRAfterUpdate(h, s->precision, true);
  else
    assembleR(h, IK_DBC_FULL);
  double f, invRoot, dfdDl;
  arcSubsidiary(h, a, f, invRoot, dfdDl);
  double rNorm        = normR(h);  // the first norm leaves f out (:244)
  info->residual_norm = rNorm;
  if (!std::isfinite(rNorm)) {
    info->success = 0;
    return;
  }
  int iter = 0;
  while (rNorm > s->tol && iter < s->max_iter) {
    // the load direction Fext0 - F_idbc(u) (:251-253)
    const double* rhsL = h->arcF.p;
    if (idbc) {
      idbcForce(h);
      h->arcRhs.ensure(n);
      IK_HIP(hipMemcpyAsync(h->arcRhs.p, h->arcF.p, n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
      ik::axpy(h, -1.0, h->idbcF.p, h->arcRhs.p, n);
      rhsL = h->arcRhs.p;
    }
    if (!kReady(h, IK_DBC_FULL, s->precision)) assembleK(h, IK_DBC_FULL, s->precision);
    ik_solver_info li{};
    linearSolve(h, h->R.p, h->arcDR.p, &s->linear, s->linear.rtol, &li);  // K x = R, dR = -x
    info->linear_iterations += li.iterations;
    hipLaunchKernelGGL(negateKernel, dim3(grid(n)), dim3(256), 0, h->stream, h->arcDR.p, n);
    linearSolve(h, rhsL, h->arcDL.p, &s->linear, s->linear.rtol, &li);  // K dL = Fext0 - F_idbc
    info->linear_iterations += li.iterations;
    const double ddR = a.predicted ? ik::dot(h, h->arcDD.p, h->arcDR.p, n) : 0.0;
    const double ddL = a.predicted ? ik::dot(h, h->arcDD.p, h->arcDL.p, n) : 0.0;
    const double den = invRoot * ddL + dfdDl;
    require(den != 0.0, IK_EINVAL,
            "path following: the subsidiary equation has no load direction (the start state is not in equilibrium)");
    const double dl = (-f - invRoot * ddR) / den;
    ik::axpy(h, dl, h->arcDL.p, h->arcDR.p, n);  // the correction dR + dl dL
    info->correction_norm = std::sqrt(ik::dot(h, h->arcDR.p, h->arcDR.p, n) + dl * dl);
    if (h->neas) ik::easUpdate(h, h->arcDR.p);  // CORRECTION_UPDATED (:268), at the pre-update state
    ik::axpy(h, 1.0, h->arcDR.p, h->u.p, n);
    ik::axpy(h, 1.0, h->arcDR.p, h->arcDD.p, n);
    h->lambda += dl;
    a.Dl += dl;
    syncIdbc(h);  // syncParameterAndGlobalSolution (:276-277)
    h->kValid        = false;
    h->easCacheValid = false;
    assembleRAfterUpdate(h, s->precision, !idbc);  // with IDBC every iteration assembles K_raw for F_idbc first
    arcSubsidiary(h, a, f, invRoot, dfdDl);
    rNorm               = std::sqrt(ik::dot(h, h->R.p, h->R.p, n) + f * f);
    info->residual_norm = rNorm;
    ++iter;
    if (!std::isfinite(rNorm)) break;
  }
  if (iter == s->max_iter || !std::isfinite(rNorm)) info->success = 0;
  info->iterations = iter;
}

}  // namespace

extern "C" {

int ik_create(ik_handle** out, const ik_problem_desc* d0) {
  if (!out || !d0) return IK_EINVAL;
  auto* h = new ik_handle;
  const int rc = guard(h, [&] {
    require(d0->dim == 3 || d0->dim == 2, IK_ENOTIMPL,
            "dim must be 3 (trilinear hex8) or 2 (bilinear Q1 in plane strain)");
    Extruded ex;
    ik_problem_desc d3;
    const ik_problem_desc* d = d0;
    if (d0->dim == 2) {
      d3 = extrude2d(h, *d0, ex);
      d  = &d3;
    }
    require(d->num_nodes > 0 && d->num_elements > 0 && d->X && d->conn, IK_EINVAL, "empty mesh");
    require(d->eas_modes == 0 || d->eas_modes == 21, IK_ENOTIMPL, "eas_modes must be 0 or 21");
    require(d->material == IK_NEOHOOKE || d->material == IK_SVK, IK_EINVAL, "unknown material");
    require(d->num_nodes * 3 < (int64_t)INT32_MAX, IK_EINVAL, "too many DOFs for int32 indices");
    // connectivity, before any device work: indices in range, eight distinct nodes per element, and every node in
    // some element (a grid vertex belongs to an element; an unreferenced node would have an empty block row)
    {
      std::vector<uint8_t> seen((size_t)d->num_nodes, 0);
      for (int64_t e = 0; e < d->num_elements; ++e) {
        const int32_t* c = d->conn + e * 8;
        for (int a = 0; a < 8; ++a) {
          require(c[a] >= 0 && c[a] < d->num_nodes, IK_EINVAL, "connectivity index out of range");
          for (int b = 0; b < a; ++b)
            require(c[a] != c[b], IK_EINVAL, "element " + std::to_string(e) + " repeats a node");
          seen[c[a]] = 1;
        }
      }
      for (int64_t v = 0; v < d->num_nodes; ++v)
        require(seen[v], IK_EINVAL, "node " + std::to_string(v) + " belongs to no element");
    }
    IK_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    IK_HIP(hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking));
    IK_HIP(hipEventCreateWithFlags(&h->joinEvent, hipEventDisableTiming));
    IK_HIP(hipEventCreate(&h->ev0));
    IK_HIP(hipEventCreate(&h->ev1));
    h->nn       = d->num_nodes;
    h->ne       = d->num_elements;
    h->ndof     = h->nn * 3;
    h->material = d->material;
    h->lam      = d->lambda;
    h->mu       = d->mu;
    h->neas     = d->eas_modes;
    h->hX.assign(d->X, d->X + h->nn * 3);
    h->hconn.assign(d->conn, d->conn + h->ne * 8);
    for (int64_t i = 0; i < h->ne * 8; ++i)
      require(h->hconn[i] >= 0 && h->hconn[i] < h->nn, IK_EINVAL, "connectivity index out of range");
    if (d->dirichlet)
      h->hfixed.assign(d->dirichlet, d->dirichlet + h->ndof);
    else
      h->hfixed.assign(h->ndof, 0);
    std::vector<int32_t> fl;
    for (int64_t i = 0; i < h->ndof; ++i)
      if (h->hfixed[i]) fl.push_back((int32_t)i);
    h->nfixed = (int64_t)fl.size();
    h->X.upload(h->hX.data(), h->hX.size(), h->stream);
    h->conn.upload(h->hconn.data(), h->hconn.size(), h->stream);
    h->fixed.upload(h->hfixed.data(), h->hfixed.size(), h->stream);
    h->fixedList.upload(fl.data(), fl.size(), h->stream);
    h->nfaces = d->num_faces;
    if (h->nfaces) {
      require(d->face_elem && d->face_axis && d->face_side, IK_EINVAL, "face arrays missing");
      for (int64_t f = 0; f < h->nfaces; ++f)
        require(d->face_elem[f] >= 0 && d->face_elem[f] < h->ne && d->face_axis[f] >= 0 && d->face_axis[f] < 3 &&
                    (d->face_side[f] == 0 || d->face_side[f] == 1),
                IK_EINVAL, "bad traction face");
      // per element, one bit per loaded face (the element residual kernels add the Traction skill)
      std::vector<uint8_t> bits(h->ne, 0);
      for (int64_t f = 0; f < h->nfaces; ++f) {
        const uint8_t b = (uint8_t)(1u << (2 * d->face_axis[f] + d->face_side[f]));
        require(!(bits[d->face_elem[f]] & b), IK_EINVAL, "traction face listed twice");
        bits[d->face_elem[f]] |= b;
      }
      h->faceBits.upload(bits.data(), bits.size(), h->stream);
    }
    for (int k = 0; k < 3; ++k) h->trac[k] = h->mlpTraction[k] = d->traction[k];
    h->npl = d->num_point_loads;
    if (h->npl) {
      require(d->point_dof && d->point_coef, IK_EINVAL, "point load arrays missing");
      for (int64_t i = 0; i < h->npl; ++i)
        require(d->point_dof[i] >= 0 && d->point_dof[i] < h->ndof, IK_EINVAL, "point load dof out of range");
      h->hplDof.assign(d->point_dof, d->point_dof + h->npl);
      h->hplCoef.assign(d->point_coef, d->point_coef + h->npl);
      h->plDof.upload(d->point_dof, h->npl, h->stream);
      h->plCoef.upload(d->point_coef, h->npl, h->stream);
    }
    h->u.ensure(h->ndof);
    h->R.ensure(h->ndof);
    ik::launchZero(h->u.p, h->ndof, h->stream);
    // error flags: [0] lowest element with det C <= 0, [1] lowest EAS element with a singular D
    h->errFlag.ensure(2);
    const unsigned long long none[2] = {~0ull, ~0ull};
    IK_HIP(hipMemcpyAsync(h->errFlag.p, none, sizeof(none), hipMemcpyHostToDevice, h->stream));
    if (h->neas) {
      h->alpha.ensure((size_t)h->ne * h->neas);
      ik::launchZero(h->alpha.p, h->ne * h->neas, h->stream);
    }
    IK_HIP(hipStreamSynchronize(h->stream));
  });
  if (rc != IK_OK) {
    // keep the handle so that ik_last_error works; caller must destroy
  }
  *out = h;
  return rc;
}

void ik_destroy(ik_handle* h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->hsc) (void)hipHostFree(h->hsc);
  if (h->hred) (void)hipHostFree(h->hred);
  if (h->herr) (void)hipHostFree(h->herr);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream2) (void)hipStreamSynchronize(h->stream2);
  ik::mlpDestroy(h->mlp);
  ik::coarseDestroy(h->coarse);
  for (auto ev : h->chunkEvents) (void)hipEventDestroy(ev);
  if (h->joinEvent) (void)hipEventDestroy(h->joinEvent);
  hipStream_t s = h->stream, s2 = h->stream2;
  delete h;
  if (s) (void)hipStreamDestroy(s);
  if (s2) (void)hipStreamDestroy(s2);
}

const char* ik_last_error(const ik_handle* h) { return h ? h->err.c_str() : "null handle"; }
int64_t ik_bad_element(const ik_handle* h) { return h ? h->bad_elem : -1; }
int64_t ik_size(const ik_handle* h) { return h ? (h->plane2d ? 2 * h->nn2 : h->ndof) : -1; }
int64_t ik_reduced_size(const ik_handle* h) {
  if (!h) return -1;
  if (!h->plane2d) return h->ndof - h->nfixed;
  return 2 * h->nn2 - (int64_t)std::count_if(h->hfixed2.begin(), h->hfixed2.end(), [](uint8_t f) { return f != 0; });
}

int ik_set_requirement(ik_handle* h, const double* u, double lambda, int where) {
  return guard(h, [&] {
    if (u && h->plane2d)
      expandIn(h, u, where, 1.0, h->u.p);
    else if (u)
      copyIn(h, h->u.p, u, h->ndof, where);
    h->lambda        = lambda;
    h->bound         = true;
    h->kValid        = false;
    h->easCacheValid = false;
    IK_HIP(hipStreamSynchronize(h->stream));
  });
}

int ik_set_dirichlet_values(ik_handle* h, const double* ubar, int where) {
  return guard(h, [&] {
    if (!ubar) {
      h->hasUbar = false;
      return;
    }
    const int64_t nIn = h->plane2d ? 2 * h->nn2 : h->ndof;
    std::vector<double> hb(nIn);
    if (where == IK_DEVICE)
      IK_HIP(hipMemcpy(hb.data(), ubar, nIn * sizeof(double), hipMemcpyDeviceToHost));
    else
      std::copy(ubar, ubar + nIn, hb.begin());
    const std::vector<uint8_t>& fl = h->plane2d ? h->hfixed2 : h->hfixed;
    for (int64_t i = 0; i < nIn; ++i)
      require(hb[i] == 0.0 || fl[i], IK_EINVAL, "inhomogeneous Dirichlet value on an unconstrained DOF");
    if (h->plane2d) {  // both layers carry the 2D values
      std::vector<double> b3(h->ndof, 0.0);
      for (int64_t v = 0; v < h->nn2; ++v)
        for (int l = 0; l < 2; ++l)
          for (int k = 0; k < 2; ++k) b3[3 * (v + l * h->nn2) + k] = hb[2 * v + k];
      hb.swap(b3);
    }
    h->ubar.upload(hb.data(), hb.size(), h->stream);
    IK_HIP(hipStreamSynchronize(h->stream));
    h->hasUbar = true;
  });
}

int ik_energy(ik_handle* h, double* energy) {
  return guard(h, [&] {
    require(energy != nullptr, IK_EINVAL, "null argument");
    requireBound(h);
    *energy = ik::energy(h);
    if (h->plane2d) *energy /= kSqrt2;  // the extrusion's thickness is sqrt(2)
  });
}

int ik_stress_at(ik_handle* h, int kind, const double xi[3], double* out, int where) {
  return guard(h, [&] {
    require(xi != nullptr && out != nullptr, IK_EINVAL, "null argument");
    requireBound(h);
    if (h->plane2d) {  // xi in [0,1]^2; plane-strain Voigt [11, 22, 12] of the 3D stress at the mid-plane
      const double x3[3] = {xi[0], xi[1], 0.5};
      h->tmp.ensure((size_t)h->ne * 6);
      ik::stressAt(h, kind, x3, h->tmp.p);
      std::vector<double> s6((size_t)h->ne * 6), s3((size_t)h->ne * 3);
      copyOut(h, s6.data(), h->tmp.p, h->ne * 6, IK_HOST);
      for (int64_t e = 0; e < h->ne; ++e) {
        s3[e * 3]     = s6[e * 6];
        s3[e * 3 + 1] = s6[e * 6 + 1];
        s3[e * 3 + 2] = s6[e * 6 + 5];
      }
      IK_HIP(hipMemcpy(out, s3.data(), s3.size() * sizeof(double),
                       where == IK_DEVICE ? hipMemcpyHostToDevice : hipMemcpyHostToHost));
      return;
    }
    if (where == IK_DEVICE) {
      ik::stressAt(h, kind, xi, out);
    } else {
      h->tmp.ensure((size_t)h->ne * 6);
      ik::stressAt(h, kind, xi, h->tmp.p);
      copyOut(h, out, h->tmp.p, h->ne * 6, IK_HOST);
    }
  });
}

int ik_get_requirement(ik_handle* h, double* u, double* lambda, int where) {
  return guard(h, [&] {
    if (u && h->plane2d)
      contractOut(h, h->u.p, 1.0, 0.0, u, where);
    else if (u)
      copyOut(h, u, h->u.p, h->ndof, where);
    if (lambda) *lambda = h->lambda;
  });
}

// copies the handle's residual out: all DOFs (RAW/FULL) or the free DOFs only (REDUCED)
static void residualOut(ik_handle* h, int dbc, double* R, int where) {
  if (!R) return;  // assembled into the handle's residual only
  if (h->plane2d) {  // each layer holds R_2D / sqrt(2)
    const double s = 1.0 / kSqrt2;
    if (dbc != IK_DBC_REDUCED) {
      contractOut(h, h->R.p, s, s, R, where);
      return;
    }
    std::vector<double> r2(2 * h->nn2), rr;
    contractOut(h, h->R.p, s, s, r2.data(), IK_HOST);
    for (int64_t i = 0; i < 2 * h->nn2; ++i)
      if (!h->hfixed2[i]) rr.push_back(r2[i]);
    if (rr.empty()) return;
    IK_HIP(hipMemcpy(R, rr.data(), rr.size() * sizeof(double),
                     where == IK_DEVICE ? hipMemcpyHostToDevice : hipMemcpyHostToHost));
    return;
  }
  if (dbc != IK_DBC_REDUCED) {
    copyOut(h, R, h->R.p, h->ndof, where);
    return;
  }
  const int64_t nr = h->ndof - h->nfixed;
  std::vector<int32_t> freeList;
  freeList.reserve(nr);
  for (int64_t i = 0; i < h->ndof; ++i)
    if (!h->hfixed[i]) freeList.push_back((int32_t)i);
  ik::DevBuf<int32_t> fl;
  fl.upload(freeList.data(), freeList.size(), h->stream);
  h->tmp.ensure(std::max<int64_t>(nr, 1));
  hipLaunchKernelGGL(gatherKernel, dim3(grid(nr)), dim3(256), 0, h->stream, h->R.p, fl.p, nr, h->tmp.p);
  copyOut(h, R, h->tmp.p, nr, where);
}

int ik_vector(ik_handle* h, int dbc, double* R, int where) {
  return guard(h, [&] {
    requireBound(h);
    require(dbc >= 0 && dbc <= 2, IK_EINVAL, "bad DBC option");
    assembleR(h, dbc == IK_DBC_FULL ? IK_DBC_FULL : IK_DBC_RAW);
    residualOut(h, dbc, R, where);
  });
}

int ik_matrix(ik_handle* h, int dbc, int precision) {
  return guard(h, [&] {
    requireBound(h);
    require(dbc >= 0 && dbc <= 2, IK_EINVAL, "bad DBC option");
    require(precision == IK_FP64 || precision == IK_FP32, IK_EINVAL, "bad precision");
    assembleK(h, dbc == IK_DBC_FULL ? IK_DBC_FULL : IK_DBC_RAW, precision);
    h->kDbc = dbc;
    IK_HIP(hipStreamSynchronize(h->stream));
  });
}

int ik_assemble(ik_handle* h, int dbc, int precision, double* R, int where) {
  return guard(h, [&] {
    requireBound(h);
    require(dbc >= 0 && dbc <= 2, IK_EINVAL, "bad DBC option");
    require(precision == IK_FP64 || precision == IK_FP32, IK_EINVAL, "bad precision");
    ensurePattern(h);
    const int d = dbc == IK_DBC_FULL ? IK_DBC_FULL : IK_DBC_RAW;
    // FP64: one element pass produces K_e and R_e; FP32: the FP32 tangent and the FP64 residual are two passes
    ik::assembleKR(h, d, precision, true, true);
    h->kPrecision = precision;
    h->kDbc       = dbc;
    h->kValid     = true;
    residualOut(h, dbc, R, where);
    IK_HIP(hipStreamSynchronize(h->stream));
  });
}

// The assembled blocks in CSR order (block b of rowptr / colidx at [9 b, 9 b + 9)), in FP64, from the device's
// sliced-ELL slots
static std::vector<double> downloadK(ik_handle* h) {
  const size_t ns = (size_t)h->nsell * 9;
  std::vector<double> dev(ns), K((size_t)h->nnzb * 9);
  if (h->kPrecision == IK_FP64) {
    IK_HIP(hipMemcpyAsync(dev.data(), h->K64.p, ns * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    IK_HIP(hipStreamSynchronize(h->stream));
  } else {
    std::vector<float> f(ns);
    IK_HIP(hipMemcpyAsync(f.data(), h->K32.p, ns * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    IK_HIP(hipStreamSynchronize(h->stream));
    for (size_t i = 0; i < ns; ++i) dev[i] = f[i];
  }
  for (int64_t b = 0; b < h->nnzb; ++b)
    std::copy_n(dev.begin() + (size_t)h->hpos[b] * 9, 9, K.begin() + (size_t)b * 9);
  return K;
}

// Eigen CSC of the plane-strain matrix from the extrusion's blocks: column 2i+k holds rows 2j+l for the 2D
// neighbours j of node i (the layer-0 columns of 3D block row i, sorted), K2 = sqrt(2) (K(j, i) + K(j, i + nn2))
// restricted to the in-plane components; FULL then sets the constrained rows and columns of the identity
// (simpleassemblers.inl:152-169), REDUCED drops them.
static void csc2d(ik_handle* h, int dbc, int64_t* n, int64_t* nnz, int32_t* outer, int32_t* inner, double* values) {
  const bool reduced = dbc == IK_DBC_REDUCED;
  const int64_t nn2 = h->nn2, nd = 2 * nn2;
  const auto& fx    = h->hfixed2;
  std::vector<int64_t> cb(nd + 1, 0);
  for (int64_t i = 0; i < nd; ++i) cb[i + 1] = cb[i] + (fx[i] ? 1 : 0);
  const int64_t N = reduced ? nd - cb[nd] : nd;
  std::vector<int64_t> colStart(N + 1, 0);
  for (int64_t i = 0; i < nn2; ++i)
    for (int k = 0; k < 2; ++k) {
      const int64_t c = 2 * i + k;
      if (reduced && fx[c]) continue;
      int64_t cnt = 0;
      for (int32_t b = h->hrowptr[i]; b < h->hrowptr[i + 1] && h->hcolidx[b] < nn2; ++b)
        for (int l = 0; l < 2; ++l) cnt += !(reduced && fx[2 * (int64_t)h->hcolidx[b] + l]);
      colStart[(reduced ? c - cb[c] : c) + 1] = cnt;
    }
  for (int64_t c = 0; c < N; ++c) colStart[c + 1] += colStart[c];
  *n   = N;
  *nnz = colStart[N];
  if (!outer && !inner && !values) return;
  std::vector<double> K;
  if (values) {
    require(h->kValid, IK_EINVAL, "no matrix assembled");
    require((dbc == IK_DBC_FULL) == (h->kDbc == IK_DBC_FULL), IK_EINVAL,
            "requested DBC layout differs from the assembled matrix");
    K = downloadK(h);
  }
  auto blockAt = [&](int64_t r, int64_t c) {  // index of block (r, c) of the 3D pattern
    const auto first = h->hcolidx.begin() + h->hrowptr[r], last = h->hcolidx.begin() + h->hrowptr[r + 1];
    const auto it    = std::lower_bound(first, last, (int32_t)c);
    require(it != last && *it == c, IK_EINVAL, "extruded pattern misses a block");
    return (int64_t)(it - h->hcolidx.begin());
  };
  if (outer)
    for (int64_t c = 0; c <= N; ++c) outer[c] = (int32_t)colStart[c];
  for (int64_t i = 0; i < nn2; ++i)
    for (int k = 0; k < 2; ++k) {
      const int64_t c = 2 * i + k;
      if (reduced && fx[c]) continue;
      int64_t p = colStart[reduced ? c - cb[c] : c];
      for (int32_t b = h->hrowptr[i]; b < h->hrowptr[i + 1] && h->hcolidx[b] < nn2; ++b) {
        const int64_t j = h->hcolidx[b];
        int64_t b0 = -1, b1 = -1;
        if (values) {
          b0 = blockAt(j, i);
          b1 = blockAt(j, i + nn2);
        }
        for (int l = 0; l < 2; ++l) {
          const int64_t r = 2 * j + l;
          if (reduced && fx[r]) continue;
          if (inner) inner[p] = (int32_t)(reduced ? r - cb[r] : r);
          if (values) {
            double v = kSqrt2 * (K[(size_t)b0 * 9 + l * 3 + k] + K[(size_t)b1 * 9 + l * 3 + k]);
            if (dbc == IK_DBC_FULL && (fx[r] || fx[c])) v = (r == c) ? 1.0 : 0.0;
            values[p] = v;
          }
          ++p;
        }
      }
    }
}

int ik_matrix_csc(ik_handle* h, int dbc, int64_t* n, int64_t* nnz, int32_t* outer, int32_t* inner, double* values) {
  return guard(h, [&] {
    ensurePattern(h);
    require(dbc >= 0 && dbc <= 2, IK_EINVAL, "bad DBC option");
    if (h->plane2d) {
      csc2d(h, dbc, n, nnz, outer, inner, values);
      return;
    }
    const bool reduced = dbc == IK_DBC_REDUCED;
    // block transpose map and per-node block counts
    const int64_t nn = h->nn;
    std::vector<int64_t> cb(h->ndof + 1, 0);  // constraints below
    for (int64_t i = 0; i < h->ndof; ++i) cb[i + 1] = cb[i] + (h->hfixed[i] ? 1 : 0);
    const int64_t N = reduced ? h->ndof - h->nfixed : h->ndof;
    // count entries per column (column c = 3*i + k of the full matrix)
    std::vector<int64_t> colStart(N + 1, 0);
    for (int64_t i = 0; i < nn; ++i)
      for (int k = 0; k < 3; ++k) {
        const int64_t c = 3 * i + k;
        if (reduced && h->hfixed[c]) continue;
        int64_t cnt = 0;
        for (int32_t b = h->hrowptr[i]; b < h->hrowptr[i + 1]; ++b)
          for (int l = 0; l < 3; ++l) {
            const int64_t r = 3 * (int64_t)h->hcolidx[b] + l;
            if (reduced && h->hfixed[r]) continue;
            ++cnt;
          }
        colStart[(reduced ? c - cb[c] : c) + 1] = cnt;
      }
    for (int64_t c = 0; c < N; ++c) colStart[c + 1] += colStart[c];
    *n   = N;
    *nnz = colStart[N];
    if (!outer && !inner && !values) return;
    std::vector<double> K;
    if (values) {
      require(h->kValid, IK_EINVAL, "no matrix assembled");
      require((dbc == IK_DBC_FULL) == (h->kDbc == IK_DBC_FULL), IK_EINVAL,
              "requested DBC layout differs from the assembled matrix");
      K = downloadK(h);
    }
    if (outer)
      for (int64_t c = 0; c <= N; ++c) outer[c] = (int32_t)colStart[c];
    for (int64_t i = 0; i < nn; ++i)
      for (int k = 0; k < 3; ++k) {
        const int64_t c = 3 * i + k;
        if (reduced && h->hfixed[c]) continue;
        int64_t p = colStart[reduced ? c - cb[c] : c];
        for (int32_t b = h->hrowptr[i]; b < h->hrowptr[i + 1]; ++b) {
          const int32_t j = h->hcolidx[b];
          // entry (row 3j+l, col 3i+k) lives in block (j, i), element (l, k)
          int64_t bt = -1;
          if (values) {
            const auto first = h->hcolidx.begin() + h->hrowptr[j], last = h->hcolidx.begin() + h->hrowptr[j + 1];
            bt = std::lower_bound(first, last, (int32_t)i) - h->hcolidx.begin();
          }
          for (int l = 0; l < 3; ++l) {
            const int64_t r = 3 * (int64_t)j + l;
            if (reduced && h->hfixed[r]) continue;
            if (inner) inner[p] = (int32_t)(reduced ? r - cb[r] : r);
            if (values) values[p] = K[(size_t)bt * 9 + l * 3 + k];
            ++p;
          }
        }
      }
  });
}

int ik_matrix_bsr(ik_handle* h, int64_t* nbr, int64_t* nnzb, const int32_t** rowptr, const int32_t** colidx,
                  const int32_t** block_pos, const void** values, int* precision) {
  return guard(h, [&] {
    require(!h->plane2d, IK_ENOTIMPL, "the block view is the 3D storage; a plane-strain handle exports CSC only");
    ensurePattern(h);
    if (block_pos && !h->blockPos.p) h->blockPos.upload(h->hpos.data(), h->hpos.size(), h->stream);
    if (colidx && !h->colidx.p) h->colidx.upload(h->hcolidx.data(), h->hcolidx.size(), h->stream);
    if (nbr) *nbr = h->nn;
    if (nnzb) *nnzb = h->nnzb;
    if (rowptr) *rowptr = h->rowptr.p;
    if (colidx) *colidx = h->colidx.p;
    if (block_pos) *block_pos = h->blockPos.p;
    if (values) *values = h->kPrecision == IK_FP32 ? (const void*)h->K32.p : (const void*)h->K64.p;
    if (precision) *precision = h->kPrecision;
    IK_HIP(hipStreamSynchronize(h->stream));
  });
}

int ik_element(ik_handle* h, int64_t e, double* Ke, double* Re) {
  return guard(h, [&] {
    requireBound(h);
    ensurePattern(h);
    if (h->plane2d) {  // the Q1 element from its hex: K2 = sqrt(2) (K(a, b) + K(a, b + 4)), R2 = (R(a) + R(a + 4)) / sqrt(2)
      std::vector<double> K3(24 * 24), R3(24);
      ik::elementHost(h, e, Ke ? K3.data() : nullptr, Re ? R3.data() : nullptr);
      for (int a = 0; a < 4; ++a)
        for (int k = 0; k < 2; ++k) {
          if (Re) Re[2 * a + k] = (R3[3 * a + k] + R3[3 * (a + 4) + k]) / kSqrt2;
          if (Ke)
            for (int b = 0; b < 4; ++b)
              for (int l = 0; l < 2; ++l)
                Ke[(2 * a + k) * 8 + 2 * b + l] =
                    kSqrt2 * (K3[(3 * a + k) * 24 + 3 * b + l] + K3[(3 * a + k) * 24 + 3 * (b + 4) + l]);
        }
      return;
    }
    ik::elementHost(h, e, Ke, Re);
  });
}

int ik_update_state(ik_handle* h, const double* du, int where) {
  return guard(h, [&] {
    requireBound(h);
    if (!h->neas) return;
    h->du.ensure(h->ndof);
    copyIn(h, h->du.p, du, h->ndof, where);
    if (!h->easCacheValid) assembleR(h, IK_DBC_RAW);  // D, L, Rt at the bound state
    ik::easUpdate(h, h->du.p);
    IK_HIP(hipStreamSynchronize(h->stream));
  });
}

int ik_get_alpha(ik_handle* h, double* alpha, int where) {
  return guard(h, [&] {
    require(h->neas > 0, IK_EINVAL, "no EAS modes");
    copyOut(h, alpha, h->alpha.p, h->ne * h->neas, where);
  });
}

int ik_set_alpha(ik_handle* h, const double* alpha, int where) {
  return guard(h, [&] {
    require(h->neas > 0, IK_EINVAL, "no EAS modes");
    copyIn(h, h->alpha.p, alpha, h->ne * h->neas, where);
    h->easCacheValid = false;
    h->kValid        = false;
    IK_HIP(hipStreamSynchronize(h->stream));
  });
}

int ik_solve(ik_handle* h, const double* rhs, double* x, int where, const ik_solver_opts* o, ik_solver_info* info) {
  return guard(h, [&] {
    require(rhs && x && o && info, IK_EINVAL, "null argument");
    const int64_t n = h->ndof;
    h->rhs.ensure(n);
    h->px.ensure(n);
    if (h->plane2d)  // K3 [x; x] = [b; b] / sqrt(2) is K2 x = b on the symmetric layer pair
      expandIn(h, rhs, where, 1.0 / kSqrt2, h->rhs.p);
    else
      copyIn(h, h->rhs.p, rhs, n, where);
    linearSolve(h, h->rhs.p, h->px.p, o, o->rtol, info);
    if (h->plane2d)
      contractOut(h, h->px.p, 1.0, 0.0, x, where);
    else
      copyOut(h, x, h->px.p, n, where);
  });
}

int ik_newton_solve(ik_handle* h, const ik_nr_settings* s, double stepSize, ik_nr_info* info) {
  return guard(h, [&] {
    require(s && info, IK_EINVAL, "null argument");
    newtonSolve(h, s, stepSize, info);
    if (h->plane2d) info->correction_norm /= kSqrt2;  // both layers carry the 2D correction
  });
}

int ik_load_control(ik_handle* h, const ik_lc_config* c, const ik_nr_settings* s, ik_control_info* info,
                    int32_t* iterations, double* residuals) {
  return guard(h, [&] {
    require(c && s && info, IK_EINVAL, "null argument");
    require(c->load_steps > 0, IK_EINVAL, "Number of load steps should be greater than zero.");
    require(!(h->plane2d && c->predictor == IK_PRED_MLP), IK_ENOTIMPL,
            "the learned predictor's features are those of a 3D mesh");
    requireBound(h);
    *info           = ik_control_info{};
    const double dl = (c->tend - c->tbegin) / c->load_steps;
    ik_nr_info ni{};
    newtonSolve(h, s, 0.0, &ni);
    if (iterations) iterations[0] = ni.iterations;
    if (residuals) residuals[0] = ni.residual_norm;
    info->solves = 1;
    info->total_iterations += ni.iterations;
    info->linear_iterations += ni.linear_iterations;
    if (!ni.success) return;
    require(c->predictor == IK_PRED_CONSTANT || c->predictor == IK_PRED_LINEAR || c->predictor == IK_PRED_MLP,
            IK_EINVAL, "unknown predictor");
    // increment history for the extrapolating predictors: duN = u_n - u_{n-1}, duN1 = u_{n-1} - u_{n-2}
    ik::DevBuf<double> uPrev, duN, duN1;
    if (c->predictor != IK_PRED_CONSTANT) {
      uPrev.ensure(h->ndof);
      duN.ensure(h->ndof);
      duN1.ensure(h->ndof);
      IK_HIP(hipMemcpyAsync(uPrev.p, h->u.p, h->ndof * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
      ik::launchZero(duN.p, h->ndof, h->stream);
      ik::launchZero(duN1.p, h->ndof, h->stream);
    }
    // the learned predictor only chooses where Newton starts: a step whose Newton solve fails from the MLP
    // prediction (det C <= 0, a singular EAS matrix, a residual that is not finite, or maxIter) is restored to
    // u_n, alpha_n, lambda_n and solved again from the secant prediction
    ik::DevBuf<double> alphaSave;
    if (c->predictor == IK_PRED_MLP && h->neas) alphaSave.ensure((size_t)h->ne * h->neas);
    for (int ls = 0; ls < c->load_steps; ++ls) {
      const double lamN = h->lambda;
      bool mlpStep      = false;
      if (c->predictor == IK_PRED_LINEAR && ls > 0) {
        ik::axpy(h, 1.0, duN.p, h->u.p, h->ndof);  // u += u_n - u_{n-1} (ArcLength secant semantics)
      } else if (c->predictor == IK_PRED_MLP) {
        if (h->neas)
          IK_HIP(hipMemcpyAsync(alphaSave.p, h->alpha.p, alphaSave.n * sizeof(double), hipMemcpyDeviceToDevice,
                                h->stream));
        ik::mlpPredict(h, duN.p, duN1.p, h->lambda + dl, dl, ls > 0 ? dl : 0.0, ls > 1 ? dl : 0.0);
        mlpStep = true;
      }
      h->lambda += dl;  // predictor: lambda += dlambda (loadcontrol.inl:55-57)
      h->kValid        = false;
      h->easCacheValid = false;
      bool failed      = false;
      if (mlpStep) {
        try {
          newtonSolve(h, s, dl, &ni);
          failed = !ni.success;
        } catch (const Error& e) {
          if (e.code != IK_ENEGDET && e.code != IK_ESINGULAR) throw;
          failed = true;
          ni     = ik_nr_info{};
        }
        if (failed) {
          info->fallbacks += 1;
          info->total_iterations += ni.iterations;
          info->linear_iterations += ni.linear_iterations;
          IK_HIP(hipMemcpyAsync(h->u.p, uPrev.p, h->ndof * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
          if (h->neas)
            IK_HIP(hipMemcpyAsync(h->alpha.p, alphaSave.p, alphaSave.n * sizeof(double), hipMemcpyDeviceToDevice,
                                  h->stream));
          if (ls > 0) ik::axpy(h, 1.0, duN.p, h->u.p, h->ndof);
          h->lambda        = lamN + dl;
          h->kValid        = false;
          h->easCacheValid = false;
        }
      }
      if (!mlpStep || failed) newtonSolve(h, s, dl, &ni);
      if (ni.iterations == 0) syncIdbc(h);  // a step that needed no correction still takes the new values
      if (iterations) iterations[ls + 1] = ni.iterations;
      if (residuals) residuals[ls + 1] = ni.residual_norm;
      info->solves += 1;
      info->total_iterations += ni.iterations;
      info->linear_iterations += ni.linear_iterations;
      if (!ni.success) return;
      if (c->predictor != IK_PRED_CONSTANT)
        hipLaunchKernelGGL(historyKernel, dim3(grid(h->ndof)), dim3(256), 0, h->stream, h->u.p, uPrev.p, duN.p,
                           duN1.p, h->ndof);
    }
    info->success = 1;
  });
}

int ik_path_following(ik_handle* h, const ik_pf_config* c, const ik_nr_settings* s, ik_control_info* info,
                      int32_t* iterations, double* residuals, double* lambdas) {
  return guard(h, [&] {
    require(c && s && info, IK_EINVAL, "null argument");
    require(c->steps > 0, IK_EINVAL, "Number of load steps should be greater than zero.");
    require(c->step_size > 0.0, IK_EINVAL, "the arc length must be positive");
    require(c->target_iterations >= 0, IK_EINVAL, "target iterations must not be negative");
    require(!h->plane2d, IK_ENOTIMPL, "path following on the device supports 3D meshes only");
    requireBound(h);
    ensurePattern(h);
    const int64_t n = h->ndof;
    *info           = ik_control_info{};
    for (auto* b : {&h->arcDD, &h->arcDR, &h->arcDL, &h->arcF}) b->ensure(n);
    const bool idbc = h->hasUbar && h->nfixed > 0;
    ArcState a;
    a.stepSize = c->step_size;
    a.ub2      = idbc ? ik::dot(h, h->ubar.p, h->ubar.p, n) : 0.0;
    ik::launchZero(h->arcDD.p, n, h->stream);
    ik_nr_info ni{};
    auto record = [&](int k) {
      if (iterations) iterations[k] = ni.iterations;
      if (residuals) residuals[k] = ni.residual_norm;
      if (lambdas) lambdas[k] = h->lambda;
      info->solves += 1;
      info->total_iterations += ni.iterations;
      info->linear_iterations += ni.linear_iterations;
    };
    arcNewton(h, s, a, &ni);  // the start state (pathfollowing.inl:40-46)
    record(0);
    if (!ni.success) return;
    // ArcLength::initialPrediction (pathfollowingfunctions.hh:131-170): K DD = -(R + F_idbc) at lambda = 1 (u as it
    // is), psi = sqrt(|DD|^2 + |ubar|^2), DD and Dl = 1 scaled to the arc length, the inhomogeneous DOFs moved to
    // Dl ubar, then psi = sqrt(|Dl ubar|^2 + |DD|^2) / Dl
    h->lambda = 1.0;
    h->kValid = false;
    if (idbc) idbcForce(h);
    assembleR(h, IK_DBC_FULL);
    if (idbc) ik::axpy(h, 1.0, h->idbcF.p, h->R.p, n);
    assembleK(h, IK_DBC_FULL, s->precision);
    {
      ik_solver_info li{};
      linearSolve(h, h->R.p, h->arcDD.p, &s->linear, s->linear.rtol, &li);
      info->linear_iterations += li.iterations;
    }
    hipLaunchKernelGGL(negateKernel, dim3(grid(n)), dim3(256), 0, h->stream, h->arcDD.p, n);
    const double dd2 = ik::dot(h, h->arcDD.p, h->arcDD.p, n);
    a.psi            = std::sqrt(dd2 + a.ub2);
    const double sc  = std::sqrt(dd2 + a.ub2 + a.psi * a.psi);
    require(sc > 0.0, IK_EINVAL, "path following: the load direction is zero");
    hipLaunchKernelGGL(scaleKernel, dim3(grid(n)), dim3(256), 0, h->stream, h->arcDD.p, a.stepSize / sc, n);
    a.Dl = a.stepSize / sc;
    ik::axpy(h, 1.0, h->arcDD.p, h->u.p, n);
    h->lambda = a.Dl;
    syncIdbc(h);
    h->kValid        = false;
    h->easCacheValid = false;
    a.psi       = std::sqrt((a.Dl * a.Dl * a.ub2 + ik::dot(h, h->arcDD.p, h->arcDD.p, n)) / (a.Dl * a.Dl));
    a.predicted = true;
    arcNewton(h, s, a, &ni);
    record(1);
    if (!ni.success) return;
    for (int ls = 1; ls < c->steps; ++ls) {
      // AdaptiveStepSizing::IterationBased (adaptivestepsizing.hh:48-90): s <- sqrt(target / its) s when the last
      // solve took more than the target iterations
      if (c->target_iterations > 0 && ni.iterations > c->target_iterations)
        a.stepSize = std::sqrt((double)c->target_iterations / ni.iterations) * a.stepSize;
      // ArcLength::intermediatePrediction (pathfollowingfunctions.hh:172-185): the last step's increments
      ik::axpy(h, 1.0, h->arcDD.p, h->u.p, n);
      h->lambda += a.Dl;
      syncIdbc(h);
      h->kValid        = false;
      h->easCacheValid = false;
      arcNewton(h, s, a, &ni);
      record(ls + 1);
      if (!ni.success) return;
    }
    info->success = 1;
  });
}

int64_t ik_mlp_num_params(void) { return ik::mlpNumParams(); }

int ik_mlp_set_params(ik_handle* h, const float* params, int where) {
  return guard(h, [&] {
    require(params != nullptr, IK_EINVAL, "null parameters");
    if (!h->mlp) h->mlp = ik::mlpCreate(h->stream);
    ik::mlpSetParams(h->mlp, params, where, h->stream);
    h->mlpLoaded = true;
  });
}

int ik_mlp_get_params(ik_handle* h, float* params, int where) {
  return guard(h, [&] {
    require(h->mlpLoaded, IK_EINVAL, "no MLP parameters set");
    ik::mlpGetParams(h->mlp, params, where, h->stream);
  });
}

int ik_mlp_set_load(ik_handle* h, const double traction[3], double charLen) {
  return guard(h, [&] {
    require(charLen > 0, IK_EINVAL, "characteristic length must be positive");
    for (int k = 0; k < 3; ++k) h->mlpTraction[k] = traction[k];
    h->mlpCharLen = charLen;
  });
}

int ik_mlp_predict(ik_handle* h, const double* du_n, const double* du_n1, double lam_next, double dlam,
                   double dlam_prev, double dlam_prev2, int where) {
  return guard(h, [&] {
    require(h->mlpLoaded, IK_EINVAL, "no MLP parameters set");
    requireBound(h);
    require(!h->plane2d, IK_ENOTIMPL, "the learned predictor works on 3D handles");
    require(du_n != nullptr && du_n1 != nullptr, IK_EINVAL, "null increment");
    const double* dn  = du_n;
    const double* dn1 = du_n1;
    ik::DevBuf<double> hn, hn1;
    if (where != IK_DEVICE) {
      hn.ensure(h->ndof);
      hn1.ensure(h->ndof);
      copyIn(h, hn.p, du_n, h->ndof, IK_HOST);
      copyIn(h, hn1.p, du_n1, h->ndof, IK_HOST);
      dn  = hn.p;
      dn1 = hn1.p;
    }
    ik::mlpPredict(h, dn, dn1, lam_next, dlam, dlam_prev, dlam_prev2);
    h->kValid        = false;
    h->easCacheValid = false;
    IK_HIP(hipStreamSynchronize(h->stream));
  });
}

int ik_timing_enable(ik_handle* h, int on) {
  return guard(h, [&] { h->timing = on != 0; });
}
int ik_timing_read(ik_handle* h, int kernel, int64_t* launches, double* totalMs) {
  return guard(h, [&] {
    require(kernel >= 0 && kernel < ik::KCOUNT, IK_EINVAL, "bad kernel id");
    if (launches) *launches = h->tLaunches[kernel];
    if (totalMs) *totalMs = h->tMs[kernel];
  });
}
int ik_timing_reset(ik_handle* h) {
  return guard(h, [&] {
    for (int i = 0; i < ik::KCOUNT; ++i) {
      h->tLaunches[i] = 0;
      h->tMs[i]       = 0;
    }
  });
}
int ik_synchronize(ik_handle* h) {
  return guard(h, [&] { IK_HIP(hipStreamSynchronize(h->stream)); });
}
int ik_assembly_layout(ik_handle* h, int64_t* staged, int64_t* gathered, int64_t* direct) {
  return guard(h, [&] {
    ensurePattern(h);
    const int64_t nup = (h->nnzb + h->nn) / 2;  // upper blocks incl. the diagonal (the pattern is symmetric)
    const int64_t ng  = h->chunkQ.empty() ? 0 : h->chunkQ.back();
    if (staged) *staged = h->nstage;
    if (gathered) *gathered = ng;
    if (direct) *direct = nup - ng;
  });
}

}  // extern "C"
