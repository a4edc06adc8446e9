// The reference's 3D cantilever test (tests/src/testcantileverbeam.hh:83-198 with the results of
// tests/src/testcantileverbeamEAS.cpp:61-67) written against the C++ mirror API of include/ikarus_amd/ikarus.hh:
// YaspGrid 10x1x1 over 10x2x2, E = 100, nu = 0.3, EAS-21 GL, x = 0 clamped, unit point loads in y at
// (10,2,2) and (10,2,0), LoadControl 20 steps over [0, 1], Newton tol 1e-10.
// Exit status 0 when 80 iterations and max|u| agree to 1e-10 (relative).
#include <cmath>
#include <cstdio>
#include <cstring>

#include "ikarus_amd/ikarus.hh"

using namespace Ikarus::AMD;

int main(int argc, char** argv) {
  const bool svk = argc > 1 && std::strcmp(argv[1], "svk") == 0;
  const int nx = 10, ny = 1, nz = 1;
  const double L = 10, h = 2;
  Problem p;
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        p.X.push_back(i * L / nx);
        p.X.push_back(j * h / ny);
        p.X.push_back(k * h / nz);
      }
  auto vid = [&](int i, int j, int k) { return i + (nx + 1) * (j + (ny + 1) * k); };
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        for (int a = 0; a < 8; ++a) p.conn.push_back(vid(i + (a & 1), j + ((a >> 1) & 1), k + ((a >> 2) & 1)));
  const double E = 100, nu = 0.3;
  p.lambda   = E * nu / ((1 + nu) * (1 - 2 * nu));
  p.mu       = E / (2 * (1 + nu));
  p.material = svk ? IK_SVK : IK_NEOHOOKE;
  p.easModes = 21;
  const int nn = (int)p.X.size() / 3;
  p.dirichlet.assign(3 * nn, 0);
  for (int v = 0; v < nn; ++v)
    if (std::fabs(p.X[3 * v]) < 1e-8)
      for (int c = 0; c < 3; ++c) p.dirichlet[3 * v + c] = 1;
  p.pointDof  = {3 * (int64_t)vid(nx, ny, nz) + 1, 3 * (int64_t)vid(nx, ny, 0) + 1};
  p.pointCoef = {1.0, 1.0};

  try {
    SparseFlatAssembler assembler(p);
    FERequirement req(std::vector<double>(3 * nn, 0.0), 0.0);
    assembler.bind(req, DBCOption::Full);
    NewtonRaphson nr(assembler, NRSettings{1e-10, 20, 0}, LinearSolverOptions{IK_PCG_JACOBI_FP64, 1e-15, 4000});
    LoadControl lc(nr, 20, {0.0, 1.0});
    const auto info = lc.run(req);
    double maxu = 0;
    for (double v : req.globalSolution()) maxu = std::max(maxu, std::fabs(v));
    const double expected = svk ? 4.7692391315649365 : 4.781820664768682;
    const bool ok = info.success && info.totalIterations == 80 && std::fabs(maxu - expected) <= 1e-10 * expected &&
                    std::fabs(req.parameter() - 1.0) < 1e-12;
    // the assembled matrix at the converged state is symmetric positive definite on the free DOFs
    const auto& K = assembler.matrix();
    const auto& R = assembler.vector();
    double rn = 0;
    for (double v : R) rn += v * v;
    std::printf("%s: iterations %d, max|u| %.16g (expected %.16g), ||R|| %.3e, K(0,0) %.6g, nnz %zu -> %s\n",
                svk ? "SVK" : "NeoHooke", info.totalIterations, maxu, expected, std::sqrt(rn), K.coeff(0, 0),
                K.inner.size(), ok ? "PASS" : "FAIL");
    return ok && std::sqrt(rn) <= 1e-10 ? 0 : 1;
  } catch (const Error& e) {
    std::printf("error %d: %s\n", e.code, e.what());
    return 2;
  }
}
