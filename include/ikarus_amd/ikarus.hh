// ikarus.hh — header-only C++ mirror of the Ikarus assembler / solver / control API over the C ABI of
// include/ikarus_amd.h. Names, argument meaning and error behaviour follow the reference:
//   SparseFlatAssembler::matrix/vector/size/reducedSize/isConstrained/constraintsBelow
//     (ikarus/assembler/simpleassemblers.hh:109, interface.hh:51-132, 355-462)
//   FERequirement (ikarus/finiteelements/ferequirements.hh:223-407)
//   NewtonRaphson / NRSettings / NonLinearSolverInformation (ikarus/solver/nonlinearsolver/newtonraphson.hh:27-32,196-257)
//   LoadControl / ControlInformation (ikarus/controlroutines/loadcontrol.hh:59-113, loadcontrol.inl:21-57)
// Errors throw Ikarus::AMD::Error (the reference throws Dune exceptions or aborts on det C <= 0).
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../ikarus_amd.h"

namespace Ikarus::AMD {

enum class DBCOption { Raw = IK_DBC_RAW, Full = IK_DBC_FULL, Reduced = IK_DBC_REDUCED };

struct Error : std::runtime_error {
  int code;
  int64_t badElement;
  Error(int c, const std::string& m, int64_t bad = -1) : std::runtime_error(m), code(c), badElement(bad) {}
};

inline void check(ik_handle* h, int rc) {
  if (rc != IK_OK) throw Error(rc, ik_last_error(h), ik_bad_element(h));
}

// FERequirement: global solution (full size) and load factor
class FERequirement {
 public:
  FERequirement(std::vector<double> u, double lambda) : u_(std::move(u)), lambda_(lambda) {}
  std::vector<double>& globalSolution() { return u_; }
  const std::vector<double>& globalSolution() const { return u_; }
  double& parameter() { return lambda_; }
  double parameter() const { return lambda_; }

 private:
  std::vector<double> u_;
  double lambda_;
};

// Eigen ColMajor CSC layout (what Eigen::SparseMatrix<double> exposes as outer/inner/values)
struct SparseMatrixCSC {
  int64_t n = 0;
  std::vector<int32_t> outer, inner;
  std::vector<double> values;
  double coeff(int64_t r, int64_t c) const {
    const auto b = inner.begin() + outer[c], e = inner.begin() + outer[c + 1];
    const auto it = std::lower_bound(b, e, (int32_t)r);
    return (it != e && *it == r) ? values[it - inner.begin()] : 0.0;
  }
};

// Mesh, material, EAS modes, Dirichlet flags and loads of one problem (owned copies)
struct Problem {
  std::vector<double> X;          // nodes x 3
  std::vector<int32_t> conn;      // elements x 8, Dune corner order
  int material = IK_NEOHOOKE;
  double lambda = 0, mu = 0;      // Lame parameters
  int easModes = 0;               // 0 or 21
  std::vector<uint8_t> dirichlet; // 3 x nodes
  std::vector<int32_t> faceElem;
  std::vector<int8_t> faceAxis, faceSide;
  std::array<double, 3> traction{0, 0, 0};
  std::vector<int64_t> pointDof;  // R[dof] += coef * lambda (AssemblerManipulator callbacks)
  std::vector<double> pointCoef;
};

class SparseFlatAssembler {
 public:
  explicit SparseFlatAssembler(const Problem& p) : fixed_(p.dirichlet) {
    ik_problem_desc d{};
    d.dim              = 3;
    d.num_nodes        = (int64_t)p.X.size() / 3;
    d.num_elements     = (int64_t)p.conn.size() / 8;
    d.X                = p.X.data();
    d.conn             = p.conn.data();
    d.material         = p.material;
    d.lambda           = p.lambda;
    d.mu               = p.mu;
    d.eas_modes        = p.easModes;
    d.dirichlet        = p.dirichlet.empty() ? nullptr : p.dirichlet.data();
    d.num_faces        = (int64_t)p.faceElem.size();
    d.face_elem        = p.faceElem.empty() ? nullptr : p.faceElem.data();
    d.face_axis        = p.faceAxis.empty() ? nullptr : p.faceAxis.data();
    d.face_side        = p.faceSide.empty() ? nullptr : p.faceSide.data();
    for (int k = 0; k < 3; ++k) d.traction[k] = p.traction[k];
    d.num_point_loads  = (int64_t)p.pointDof.size();
    d.point_dof        = p.pointDof.empty() ? nullptr : p.pointDof.data();
    d.point_coef       = p.pointCoef.empty() ? nullptr : p.pointCoef.data();
    ik_handle* h       = nullptr;
    const int rc       = ik_create(&h, &d);
    h_.reset(h);
    check(h, rc);
    if (fixed_.empty()) fixed_.assign(size(), 0);
    constraintsBelow_.assign(size(), 0);
    for (int64_t i = 1; i < size(); ++i) constraintsBelow_[i] = constraintsBelow_[i - 1] + (fixed_[i - 1] ? 1 : 0);
  }

  void bind(FERequirement& req, DBCOption dbc = DBCOption::Full) {
    req_ = &req;
    dbc_ = dbc;
  }

  int64_t size() const { return ik_size(h_.get()); }
  int64_t reducedSize() const { return ik_reduced_size(h_.get()); }
  bool isConstrained(int64_t i) const { return fixed_[i] != 0; }
  int64_t constraintsBelow(int64_t i) const { return constraintsBelow_[i]; }
  DBCOption dBCOption() const { return dbc_; }

  std::vector<double> createFullVector(const std::vector<double>& red) const {
    std::vector<double> full(size(), 0.0);
    for (int64_t i = 0, r = 0; i < size(); ++i)
      if (!fixed_[i]) full[i] = red[r++];
    return full;
  }
  std::vector<double> createReducedVector(const std::vector<double>& full) const {
    std::vector<double> red;
    red.reserve(reducedSize());
    for (int64_t i = 0; i < size(); ++i)
      if (!fixed_[i]) red.push_back(full[i]);
    return red;
  }

  // MatrixAssembler::matrix(req, stiffness, dbc): assembler-owned, valid until the next call
  const SparseMatrixCSC& matrix(const FERequirement& req, DBCOption dbc) {
    sync(req);
    check(h_.get(), ik_matrix(h_.get(), (int)dbc, IK_FP64));
    int64_t n = 0, nnz = 0;
    check(h_.get(), ik_matrix_csc(h_.get(), (int)dbc, &n, &nnz, nullptr, nullptr, nullptr));
    mat_.n = n;
    mat_.outer.resize(n + 1);
    mat_.inner.resize(nnz);
    mat_.values.resize(nnz);
    check(h_.get(), ik_matrix_csc(h_.get(), (int)dbc, &n, &nnz, mat_.outer.data(), mat_.inner.data(),
                                  mat_.values.data()));
    return mat_;
  }
  const SparseMatrixCSC& matrix() { return matrix(bound(), dbc_); }

  // VectorAssembler::vector(req, forces, dbc)
  const std::vector<double>& vector(const FERequirement& req, DBCOption dbc) {
    sync(req);
    vec_.assign(dbc == DBCOption::Reduced ? reducedSize() : size(), 0.0);
    check(h_.get(), ik_vector(h_.get(), (int)dbc, vec_.data(), IK_HOST));
    return vec_;
  }
  const std::vector<double>& vector() { return vector(bound(), dbc_); }

  ik_handle* handle() const { return h_.get(); }

 private:
  struct Deleter {
    void operator()(ik_handle* h) const { ik_destroy(h); }
  };
  std::unique_ptr<ik_handle, Deleter> h_;
  std::vector<uint8_t> fixed_;
  std::vector<int64_t> constraintsBelow_;
  FERequirement* req_ = nullptr;
  DBCOption dbc_      = DBCOption::Full;
  SparseMatrixCSC mat_;
  std::vector<double> vec_;

  const FERequirement& bound() const {
    if (!req_) throw Error(IK_ENOTBOUND, "The assembler is not bound to a requirement, affordance or dBCOption.");
    return *req_;
  }
  void sync(const FERequirement& req) {
    check(h_.get(), ik_set_requirement(h_.get(), req.globalSolution().data(), req.parameter(), IK_HOST));
  }
};

struct NRSettings {
  double tol  = 1e-8;
  int maxIter = 20;
  int minIter = 0;
};

// linear solver choice for the Newton iteration (the reference's si_ConjugateGradient is the FP64 form)
struct LinearSolverOptions {
  int kind         = IK_PCG_JACOBI_FP64;
  double rtol      = 1e-10;
  int maxIter      = 0;
  double innerRtol = 1e-3;
  int maxRefine    = 20;
  int precision    = IK_FP64;
  double forcing   = 0.0;
};

struct NonLinearSolverInformation {
  bool success = false;
  int iterations = 0;
  double residualNorm = 0, correctionNorm = 0;
  int64_t linearIterations = 0;
};

class NewtonRaphson {
 public:
  NewtonRaphson(SparseFlatAssembler& a, NRSettings s = {}, LinearSolverOptions l = {}) : a_(a), s_(s), l_(l) {}

  NonLinearSolverInformation solve(FERequirement& req, double stepSize = 0.0) {
    ik_handle* h = a_.handle();
    check(h, ik_set_requirement(h, req.globalSolution().data(), req.parameter(), IK_HOST));
    const ik_nr_settings st = settings();
    ik_nr_info info{};
    check(h, ik_newton_solve(h, &st, stepSize, &info));
    check(h, ik_get_requirement(h, req.globalSolution().data(), &req.parameter(), IK_HOST));
    return {info.success != 0, info.iterations, info.residual_norm, info.correction_norm, info.linear_iterations};
  }

  ik_nr_settings settings() const {
    ik_nr_settings st{};
    st.tol               = s_.tol;
    st.max_iter          = s_.maxIter;
    st.min_iter          = s_.minIter;
    st.precision         = l_.precision;
    st.linear.kind       = l_.kind;
    st.linear.rtol       = l_.rtol;
    st.linear.max_iter   = l_.maxIter;
    st.linear.inner_rtol = l_.innerRtol;
    st.linear.max_refine = l_.maxRefine;
    st.forcing           = l_.forcing;
    return st;
  }
  SparseFlatAssembler& assembler() { return a_; }

 private:
  SparseFlatAssembler& a_;
  NRSettings s_;
  LinearSolverOptions l_;
};

struct ControlInformation {
  bool success = false;
  int totalIterations = 0;
  int fallbacks       = 0;      // IK_PRED_MLP steps solved again from the secant prediction
  std::vector<int> iterations;  // per Newton solve, the initial one first
  std::vector<double> residualNorms;
};

class LoadControl {
 public:
  LoadControl(NewtonRaphson& nr, int loadSteps, std::array<double, 2> tbeginEnd, int predictor = IK_PRED_CONSTANT)
      : nr_(nr), steps_(loadSteps), t_(tbeginEnd), predictor_(predictor) {
    if (loadSteps <= 0) throw Error(IK_EINVAL, "Number of load steps should be greater than zero.");
  }

  ControlInformation run(FERequirement& req) {
    ik_handle* h = nr_.assembler().handle();
    check(h, ik_set_requirement(h, req.globalSolution().data(), req.parameter(), IK_HOST));
    const ik_nr_settings st = nr_.settings();
    ik_lc_config c{steps_, t_[0], t_[1], predictor_};
    ik_control_info info{};
    std::vector<int32_t> it(steps_ + 1, 0);
    std::vector<double> res(steps_ + 1, 0.0);
    check(h, ik_load_control(h, &c, &st, &info, it.data(), res.data()));
    check(h, ik_get_requirement(h, req.globalSolution().data(), &req.parameter(), IK_HOST));
    ControlInformation ci;
    ci.success         = info.success != 0;
    ci.totalIterations = info.total_iterations;
    ci.fallbacks       = info.fallbacks;
    ci.iterations.assign(it.begin(), it.begin() + info.solves);
    ci.residualNorms.assign(res.begin(), res.begin() + info.solves);
    return ci;
  }

 private:
  NewtonRaphson& nr_;
  int steps_;
  std::array<double, 2> t_;
  int predictor_;
};

// PathFollowing(nr, loadSteps, stepSize) with the ArcLength technique (controlroutines/pathfollowing.hh);
// the solve count is loadSteps + 1 (the start-state solve first), the load factor after each solve in
// loadFactors
class PathFollowing {
 public:
  // targetIterations > 0: AdaptiveStepSizing::IterationBased with that target (adaptivestepsizing.hh:48-90)
  PathFollowing(NewtonRaphson& nr, int loadSteps, double stepSize, int targetIterations = 0)
      : nr_(nr), steps_(loadSteps), s_(stepSize), target_(targetIterations) {
    if (loadSteps <= 0) throw Error(IK_EINVAL, "Number of load steps should be greater than zero.");
  }

  ControlInformation run(FERequirement& req) {
    ik_handle* h = nr_.assembler().handle();
    check(h, ik_set_requirement(h, req.globalSolution().data(), req.parameter(), IK_HOST));
    const ik_nr_settings st = nr_.settings();
    ik_pf_config c{steps_, s_, target_};
    ik_control_info info{};
    std::vector<int32_t> it(steps_ + 1, 0);
    std::vector<double> res(steps_ + 1, 0.0);
    loadFactors.assign(steps_ + 1, 0.0);
    check(h, ik_path_following(h, &c, &st, &info, it.data(), res.data(), loadFactors.data()));
    check(h, ik_get_requirement(h, req.globalSolution().data(), &req.parameter(), IK_HOST));
    ControlInformation ci;
    ci.success         = info.success != 0;
    ci.totalIterations = info.total_iterations;
    ci.iterations.assign(it.begin(), it.begin() + info.solves);
    ci.residualNorms.assign(res.begin(), res.begin() + info.solves);
    loadFactors.resize(info.solves);
    return ci;
  }

  std::vector<double> loadFactors;

 private:
  NewtonRaphson& nr_;
  int steps_;
  double s_;
  int target_;
};

}  // namespace Ikarus::AMD
