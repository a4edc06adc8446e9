// capi.cpp — plain-C entry points of the CPU oracle, loaded by tests/ and bench.py's cpu_baseline leg
// through ctypes. TEST INFRASTRUCTURE ONLY: the product library never links or calls this.
#include <cstdio>
#include <exception>
#include <string>

#include "ikoracle.hh"

using namespace ikoracle;

namespace {
thread_local std::string g_err;
struct Handle {
  Problem P;
  Solver* S = nullptr;
  ~Handle() { delete S; }
};
template <class F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const MaterialError& e) {
    g_err = e.what();
    return -2;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}
Solver& solver(Handle* h) {
  if (!h->S) h->S = new Solver(&h->P);
  return *h->S;
}
}  // namespace

extern "C" {

const char* ikor_last_error() { return g_err.c_str(); }

void ikor_lame(double E, double nu, double* lambda, double* mu) { lameFromEnu(E, nu, *lambda, *mu); }

// Structured mesh with seeded interior jitter. X: nn*dim, conn: ne*2^dim (caller sized).
int ikor_structured_mesh(int dim, const int* nel, const double* L, double jitter, uint64_t seed, double* X,
                         int* conn) {
  return guard([&] {
    Problem P;
    structuredMesh(dim, nel, L, jitter, seed, P);
    std::copy(P.X.begin(), P.X.end(), X);
    std::copy(P.conn.begin(), P.conn.end(), conn);
  });
}

void* ikor_create(int dim, int nn, int ne, const double* X, const int* conn, int matKind, double lambda, double mu,
                  int neas, int nq1d) {
  auto* h     = new Handle;
  Problem& P  = h->P;
  P.dim       = dim;
  P.nn        = nn;
  P.ne        = ne;
  P.X.assign(X, X + (size_t)nn * dim);
  P.conn.assign(conn, conn + (size_t)ne * (1 << dim));
  P.mat.kind   = matKind;
  P.mat.lambda = lambda;
  P.mat.mu     = mu;
  P.neas       = neas;
  P.nq1d       = nq1d;
  P.fixed.assign((size_t)nn * dim, 0);
  P.alpha.assign((size_t)ne * neas, 0.0);
  return h;
}

void ikor_destroy(void* hp) { delete static_cast<Handle*>(hp); }

void ikor_set_fixed(void* hp, const uint8_t* flags) {
  auto* h = static_cast<Handle*>(hp);
  h->P.fixed.assign(flags, flags + h->P.ndof());
  delete h->S;
  h->S = nullptr;
}

void ikor_set_ubar(void* hp, const double* ubar) {
  auto* h = static_cast<Handle*>(hp);
  if (ubar)
    h->P.ubar.assign(ubar, ubar + h->P.ndof());
  else
    h->P.ubar.clear();
}

void ikor_set_traction(void* hp, int nfaces, const int* elem, const int* axis, const int* side, const double* t0) {
  auto* h = static_cast<Handle*>(hp);
  h->P.faces.clear();
  for (int i = 0; i < nfaces; ++i) h->P.faces.push_back({elem[i], axis[i], side[i]});
  for (int k = 0; k < 3; ++k) h->P.t0[k] = t0 ? t0[k] : 0.0;
}

void ikor_add_point_load(void* hp, int64_t dof, double coef) {
  static_cast<Handle*>(hp)->P.pointLoads.emplace_back((int)dof, coef);
}

void ikor_set_faithful(void* hp, int on) { static_cast<Handle*>(hp)->P.faithfulCost = on != 0; }
// element loops of matrix()/vector() on n threads over an element coloring (timing variant; 1 = serial)
void ikor_set_threads(void* hp, int n) { solver(static_cast<Handle*>(hp)).asmb.threads = n < 1 ? 1 : n; }

void ikor_set_alpha(void* hp, const double* a) {
  auto* h = static_cast<Handle*>(hp);
  std::copy(a, a + h->P.alpha.size(), h->P.alpha.begin());
}
void ikor_get_alpha(void* hp, double* a) {
  auto* h = static_cast<Handle*>(hp);
  std::copy(h->P.alpha.begin(), h->P.alpha.end(), a);
}

int64_t ikor_pattern_nnz(void* hp, int reduced) {
  auto* h = static_cast<Handle*>(hp);
  int64_t nnz = -1;
  guard([&] { nnz = (int64_t)solver(h).asmb.pattern(reduced != 0).inner.size(); });
  return nnz;
}

int64_t ikor_pattern_n(void* hp, int reduced) {
  auto* h = static_cast<Handle*>(hp);
  int64_t n = -1;
  guard([&] { n = solver(h).asmb.pattern(reduced != 0).n; });
  return n;
}

int ikor_pattern(void* hp, int reduced, int32_t* outer, int32_t* inner) {
  auto* h = static_cast<Handle*>(hp);
  return guard([&] {
    const CSC& m = solver(h).asmb.pattern(reduced != 0);
    std::copy(m.outer.begin(), m.outer.end(), outer);
    std::copy(m.inner.begin(), m.inner.end(), inner);
  });
}

// element slot map of the full pattern: (nd)^2 entries, column-major local order
int ikor_slots(void* hp, int e, int64_t* out) {
  auto* h = static_cast<Handle*>(hp);
  return guard([&] {
    const CSC& m = solver(h).asmb.pattern(false);
    std::copy(m.slots[e].begin(), m.slots[e].end(), out);
  });
}

int ikor_matrix(void* hp, const double* u, double lambda, int dbc, double* values) {
  auto* h = static_cast<Handle*>(hp);
  return guard([&] {
    auto& v = solver(h).asmb.matrix(u, lambda, dbc);
    std::copy(v.begin(), v.end(), values);
  });
}

int ikor_vector(void* hp, const double* u, double lambda, int dbc, double* R) {
  auto* h = static_cast<Handle*>(hp);
  return guard([&] {
    auto v = solver(h).asmb.vector(u, lambda, dbc);
    std::copy(v.begin(), v.end(), R);
  });
}

// stresses at the local point xi of every element (3D plain elements), out[ne * 6]
int ikor_stress_at(void* hp, const double* u, const double* xi, int kind, double* out) {
  auto* h = static_cast<Handle*>(hp);
  return guard([&] {
    if (h->P.dim != 3 || h->P.neas != 0) throw std::runtime_error("stress_at: 3D plain elements only");
    for (int e = 0; e < h->P.ne; ++e) stressAt3(h->P, e, u, xi, kind, out + (size_t)e * 6);
  });
}

int ikor_energy(void* hp, const double* u, double lambda, double* energy) {
  auto* h = static_cast<Handle*>(hp);
  return guard([&] { *energy = solver(h).asmb.energy(u, lambda); });
}

// calculateMatrix / calculateVector for one element (Ke row-major nd x nd, Re nd). Optional EAS blocks.
int ikor_element(void* hp, int e, const double* u, double lambda, double* Ke, double* Re, double* Dm, double* Lm,
                 double* Rt) {
  auto* h = static_cast<Handle*>(hp);
  return guard([&] {
    ElementOut eo;
    elementComputeDispatch(h->P, e, u, lambda, true, true, eo);
    const int nd = h->P.nodesPerElem() * h->P.dim, ne = h->P.neas;
    if (Ke) std::copy(eo.K, eo.K + nd * nd, Ke);
    if (Re) std::copy(eo.R, eo.R + nd, Re);
    if (ne && Dm) std::copy(eo.Dm, eo.Dm + ne * ne, Dm);
    if (ne && Lm) std::copy(eo.Lm, eo.Lm + ne * nd, Lm);
    if (ne && Rt) std::copy(eo.Rt, eo.Rt + ne, Rt);
  });
}

int ikor_update_state(void* hp, const double* u, double lambda, const double* du) {
  auto* h = static_cast<Handle*>(hp);
  return guard([&] {
    std::vector<double> uu(u, u + h->P.ndof()), dd(du, du + h->P.ndof());
    solver(h).alphaUpdate(uu, lambda, dd);
  });
}

// One Jacobi-CG solve (Eigen semantics) on the Full-DBC matrix at (u, lambda).
int ikor_cg(void* hp, const double* u, double lambda, const double* b, double* x, double tol, int maxIter,
            int threads, int* iters, double* err) {
  auto* h = static_cast<Handle*>(hp);
  return guard([&] {
    Solver& S    = solver(h);
    auto& K      = S.asmb.matrix(u, lambda, Full);
    const CSC& m = S.asmb.pattern(false);
    auto ci      = jacobiCG(m, K, b, x, tol, maxIter > 0 ? maxIter : (int)(2 * m.n), threads);
    *iters       = ci.iterations;
    *err         = ci.error;
  });
}

// Direct (banded LDL^T) solve on the Full-DBC matrix at (u, lambda).
int ikor_direct_solve(void* hp, const double* u, double lambda, const double* b, double* x) {
  auto* h = static_cast<Handle*>(hp);
  return guard([&] {
    Solver& S    = solver(h);
    auto& K      = S.asmb.matrix(u, lambda, Full);
    const CSC& m = S.asmb.pattern(false);
    BandedLDLT f;
    if (!f.factorize(m, K)) throw std::runtime_error("LDLT factorization failed");
    f.solve(b, x);
  });
}

// LoadControl::run with NewtonRaphson. iters/res: arrays of steps+1 entries (initial solve first).
int ikor_load_control(void* hp, int steps, double tbegin, double tend, double tol, int maxIter, int linSolver,
                      double cgTol, int cgMaxIter, int threads, int predictor, double* u, double* lambda, int* iters,
                      double* res, int* total, int64_t* cgIters, int* success) {
  auto* h = static_cast<Handle*>(hp);
  return guard([&] {
    Solver& S      = solver(h);
    S.nr.tol       = tol;
    S.nr.maxIter   = maxIter;
    S.nr.linSolver = linSolver;
    S.nr.cgTol     = cgTol;
    S.nr.cgMaxIter = cgMaxIter;
    S.nr.threads   = threads;
    S.asmb.threads = threads;
    std::vector<double> uu(u, u + h->P.ndof());
    double lam = *lambda;
    auto ci    = loadControl(S, uu, lam, steps, tbegin, tend, predictor);
    std::copy(uu.begin(), uu.end(), u);
    *lambda = lam;
    for (size_t i = 0; i < ci.iterations.size(); ++i) {
      if (iters) iters[i] = ci.iterations[i];
      if (res) res[i] = ci.residualNorms[i];
    }
    *total   = ci.totalIterations;
    *cgIters = ci.cgIterations;
    *success = ci.success ? 1 : 0;
  });
}

// Single NewtonRaphson::solve at the given (u, lambda) with load increment stepSize (IDBC force scale).
int ikor_newton(void* hp, double* u, double lambda, double stepSize, double tol, int maxIter, int linSolver,
                double cgTol, int threads, int* iters, double* res, int* success) {
  auto* h = static_cast<Handle*>(hp);
  return guard([&] {
    Solver& S      = solver(h);
    S.nr.tol       = tol;
    S.nr.maxIter   = maxIter;
    S.nr.linSolver = linSolver;
    S.nr.cgTol     = cgTol;
    S.nr.threads   = threads;
    S.asmb.threads = threads;
    std::vector<double> uu(u, u + h->P.ndof());
    double lam = lambda;
    auto info  = S.solve(uu, lam, stepSize);
    std::copy(uu.begin(), uu.end(), u);
    *iters   = info.iterations;
    *res     = info.residualNorm;
    *success = info.success ? 1 : 0;
  });
}
}
