// ikoracle.hh — FP64 CPU restatement of Ikarus's nonlinear-mechanics hot path.
//
// TEST INFRASTRUCTURE ONLY. This code is the checker for the HIP library: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and never as the thing
// measured or shipped. The product library (ikarus_amd/) neither links nor calls it.
//
// Every routine cites the reference file:line it restates (paths relative to the Ikarus v0.4.9
// source tree). Pinning (what fixes each routine independently of this restatement):
//   * element kinematics / B-operator / geometric stiffness / NH / SVK / plane strain
//       - single-element |eig(K)| golden values (tests/src/testnonlinearelasticity.hh:248-338)
//       - cantilever and elastic-strip end-to-end golden values (testcantileverbeamEAS.cpp:29,65,67;
//         testinhomogeneousdbc.cpp:29-31)
//       - K = dR/du and R = dPi/du by finite differences, K(u=0) = linear K, rigid-body null space
//   * EAS E4 / E21 condensation + alpha update: cantilever golden values (testcantileverbeamEAS.cpp:29,65,67),
//       EAS with alpha=0 and no modes reduces to the displacement element, orthogonality int M detJ = 0
//   * CSC pattern / slot maps: brute-force dense comparison; Eigen setFromTriplets semantics
//   * Newton / LoadControl semantics: iteration counts of the golden tests above
//   * Jacobi-CG: Eigen ConjugateGradient algorithm (linearsolver.cpp:23-24); checked against the
//       banded direct solve
//   * banded LDL^T: checked against dense Gaussian elimination
//
// Third-party arithmetic restated here without a vendored source: dune-localfefunctions'
// greenLagrangeStrains (value, first and second derivative; spec docs/website/01_framework/
// localFunctions.md:395) and dune-geometry's Gauss-Legendre rule / multilinear geometry.
#pragma once
#include <algorithm>
#include <cmath>
#include <exception>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace ikoracle {

// ----------------------------------------------------------------------------------------------
// Small dense helpers
// ----------------------------------------------------------------------------------------------
inline double det3(const double A[3][3]) {
  return A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
         A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
}
inline void inv3(const double A[3][3], double I[3][3], double* detOut = nullptr) {
  const double d = det3(A);
  if (detOut) *detOut = d;
  const double id = 1.0 / d;
  I[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) * id;
  I[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
  I[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
  I[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) * id;
  I[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
  I[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
  I[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) * id;
  I[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
  I[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
}

// Gauss-Jordan inverse with partial pivoting (n <= 32), in place on a row-major n x n array.
inline bool invertDense(int n, double* A) {
  std::vector<double> M(A, A + n * n), I(n * n, 0.0);
  for (int i = 0; i < n; ++i) I[i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(M[r * n + c]) > std::fabs(M[p * n + c])) p = r;
    if (M[p * n + c] == 0.0) return false;
    if (p != c)
      for (int k = 0; k < n; ++k) {
        std::swap(M[p * n + k], M[c * n + k]);
        std::swap(I[p * n + k], I[c * n + k]);
      }
    const double piv = 1.0 / M[c * n + c];
    for (int k = 0; k < n; ++k) {
      M[c * n + k] *= piv;
      I[c * n + k] *= piv;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = M[r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        M[r * n + k] -= f * M[c * n + k];
        I[r * n + k] -= f * I[c * n + k];
      }
    }
  }
  std::copy(I.begin(), I.end(), A);
  return true;
}

// Voigt index pairs, order [11,22,33,23,13,12] in 3D and [11,22,12] in 2D
// (ikarus/utils/tensorutils.hh:181-202).
static const int kVoigt3[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};
static const int kVoigt2[3][2] = {{0, 0}, {1, 1}, {0, 1}};
inline int voigt3(int i, int j) {
  if (i == j) return i;
  if (i + j == 3) return 3;  // (1,2)
  if (i + j == 2) return 4;  // (0,2)
  return 5;                  // (0,1)
}

// ----------------------------------------------------------------------------------------------
// Materials (3D, Green-Lagrange input in Voigt with engineering shear)
// ----------------------------------------------------------------------------------------------
enum MaterialKind { NeoHooke = 0, StVenantKirchhoff = 1 };

struct Material {
  int kind = NeoHooke;
  double lambda = 0, mu = 0;
};

// Lame conversion (ikarus/finiteelements/physicshelper.hh:276-281 via ConvertLameConstants).
inline void lameFromEnu(double E, double nu, double& lambda, double& mu) {
  lambda = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  mu     = E / (2.0 * (1.0 + nu));
}

struct MaterialError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// S (Voigt, 6), C (Voigt 6x6 row-major), psi from E (Voigt, 6).
// NeoHooke: neohooke.hh:79-142 with C = 2E + I (strainconversions.hh:119-134);
// SVK: svk.hh:78-164. Voigt moduli: tensorutils.hh:219-227 (no extra factors).
inline void material3d(const Material& m, const double Ev[6], double S[6], double* Cv, double* psi) {
  if (m.kind == StVenantKirchhoff) {
    const double tr = Ev[0] + Ev[1] + Ev[2];
    for (int i = 0; i < 3; ++i) S[i] = m.lambda * tr + 2.0 * m.mu * Ev[i];
    for (int i = 3; i < 6; ++i) S[i] = m.mu * Ev[i];
    if (Cv) {
      for (int i = 0; i < 36; ++i) Cv[i] = 0.0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Cv[i * 6 + j] = m.lambda + (i == j ? 2.0 * m.mu : 0.0);
      for (int i = 3; i < 6; ++i) Cv[i * 6 + i] = m.mu;
    }
    if (psi) {
      const double sq = Ev[0] * Ev[0] + Ev[1] * Ev[1] + Ev[2] * Ev[2] +
                        0.5 * (Ev[3] * Ev[3] + Ev[4] * Ev[4] + Ev[5] * Ev[5]);
      *psi = 0.5 * m.lambda * tr * tr + m.mu * sq;
    }
    return;
  }
  // Neo-Hooke in terms of the right Cauchy-Green tensor
  double C[3][3];
  for (int v = 0; v < 6; ++v) {
    const int i = kVoigt3[v][0], j = kVoigt3[v][1];
    const double e = (v < 3) ? Ev[v] : 0.5 * Ev[v];
    C[i][j] = C[j][i] = 2.0 * e + (i == j ? 1.0 : 0.0);
  }
  double Ci[3][3], detC;
  inv3(C, Ci, &detC);
  if (!(detC > 0.0)) throw MaterialError("Determinant of right Cauchy Green tensor C must be greater than zero");
  const double lnJ = std::log(std::sqrt(detC));
  for (int v = 0; v < 6; ++v) {
    const int i = kVoigt3[v][0], j = kVoigt3[v][1];
    S[v] = m.mu * ((i == j ? 1.0 : 0.0) - Ci[i][j]) + m.lambda * lnJ * Ci[i][j];
  }
  if (Cv) {
    const double c2 = m.mu - m.lambda * lnJ;
    for (int p = 0; p < 6; ++p)
      for (int q = 0; q < 6; ++q) {
        const int I = kVoigt3[p][0], J = kVoigt3[p][1], K = kVoigt3[q][0], L = kVoigt3[q][1];
        Cv[p * 6 + q] = m.lambda * Ci[I][J] * Ci[K][L] + c2 * (Ci[I][K] * Ci[J][L] + Ci[I][L] * Ci[J][K]);
      }
  }
  if (psi) {
    const double trC = C[0][0] + C[1][1] + C[2][2];
    *psi = 0.5 * m.mu * (trC - 3.0 - 2.0 * lnJ) + 0.5 * m.lambda * lnJ * lnJ;
  }
}

// Plane strain: E33 = E13 = E23 = 0, keep Voigt {0,1,5}
// (ikarus/finiteelements/mechanics/materials/vanishingstrain.hh:93-120,148-168,196-198).
inline void material2dPlaneStrain(const Material& m, const double Ev2[3], double S2[3], double* Cv2, double* psi) {
  const double Ev[6] = {Ev2[0], Ev2[1], 0.0, 0.0, 0.0, Ev2[2]};
  double S[6], Cv[36];
  material3d(m, Ev, S, Cv2 ? Cv : nullptr, psi);
  static const int keep[3] = {0, 1, 5};
  for (int i = 0; i < 3; ++i) S2[i] = S[keep[i]];
  if (Cv2)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Cv2[i * 3 + j] = Cv[keep[i] * 6 + keep[j]];
}

template <int D>
inline void materialD(const Material& m, const double* Ev, double* S, double* Cv, double* psi) {
  if constexpr (D == 3)
    material3d(m, Ev, S, Cv, psi);
  else
    material2dPlaneStrain(m, Ev, S, Cv, psi);
}

// ----------------------------------------------------------------------------------------------
// Reference element, Q1 shape functions, Gauss rule
// ----------------------------------------------------------------------------------------------
// Corner a of [0,1]^D has coordinates (a&1, a>>1&1, a>>2&1): Dune reference cube numbering.
template <int D>
inline void q1Shape(const double xi[D], double N[1 << D], double dN[1 << D][D]) {
  constexpr int nn = 1 << D;
  for (int a = 0; a < nn; ++a) {
    double f[D], df[D];
    for (int k = 0; k < D; ++k) {
      const bool hi = (a >> k) & 1;
      f[k]  = hi ? xi[k] : 1.0 - xi[k];
      df[k] = hi ? 1.0 : -1.0;
    }
    double n = 1.0;
    for (int k = 0; k < D; ++k) n *= f[k];
    N[a] = n;
    for (int j = 0; j < D; ++j) {
      double d = df[j];
      for (int k = 0; k < D; ++k)
        if (k != j) d *= f[k];
      dN[a][j] = d;
    }
  }
}

// Gauss-Legendre rule on [0,1] with npts in {1,2}: the rule defaultQuadratureRule(element, 2*order)
// selects for Q1 (ikarus/utils/quadraturerulehelper.hh:62-71; nonlinearelastic.hh:124).
inline void gauss1d(int npts, double* x, double* w) {
  if (npts == 1) {
    x[0] = 0.5;
    w[0] = 1.0;
  } else if (npts == 2) {
    const double d = 0.5 / std::sqrt(3.0);
    x[0] = 0.5 - d;
    x[1] = 0.5 + d;
    w[0] = w[1] = 0.5;
  } else {
    throw std::invalid_argument("gauss1d: only 1 or 2 points supported");
  }
}

template <int D>
struct QuadPoint {
  double xi[D];
  double w;
};
template <int D>
inline std::vector<QuadPoint<D>> tensorGauss(int npts) {
  double x[2], w[2];
  gauss1d(npts, x, w);
  std::vector<QuadPoint<D>> r;
  int total = 1;
  for (int k = 0; k < D; ++k) total *= npts;
  for (int q = 0; q < total; ++q) {
    QuadPoint<D> p;
    p.w   = 1.0;
    int t = q;
    for (int k = 0; k < D; ++k) {
      const int i = t % npts;
      t /= npts;
      p.xi[k] = x[i];
      p.w *= w[i];
    }
    r.push_back(p);
  }
  return r;
}

// ----------------------------------------------------------------------------------------------
// Geometry at a point: J_ij = dx_i/dxi_j, g_a = J^{-T} dN_a/dxi (grad w.r.t. reference config)
// ----------------------------------------------------------------------------------------------
template <int D>
struct GeoPoint {
  double N[1 << D];
  double g[1 << D][D];  // dN_a/dX
  double J[D][D];       // dx_i/dxi_j
  double detJ;
};

template <int D>
inline void geoAt(const double* Xe, const double xi[D], GeoPoint<D>& gp) {
  constexpr int nn = 1 << D;
  double dN[nn][D];
  q1Shape<D>(xi, gp.N, dN);
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) {
      double s = 0.0;
      for (int a = 0; a < nn; ++a) s += Xe[a * D + i] * dN[a][j];
      gp.J[i][j] = s;
    }
  double Ji[D][D];
  if constexpr (D == 3) {
    inv3(gp.J, Ji, &gp.detJ);
  } else {
    gp.detJ = gp.J[0][0] * gp.J[1][1] - gp.J[0][1] * gp.J[1][0];
    const double id = 1.0 / gp.detJ;
    Ji[0][0] = gp.J[1][1] * id;
    Ji[0][1] = -gp.J[0][1] * id;
    Ji[1][0] = -gp.J[1][0] * id;
    Ji[1][1] = gp.J[0][0] * id;
  }
  for (int a = 0; a < nn; ++a)
    for (int i = 0; i < D; ++i) {
      double s = 0.0;
      for (int j = 0; j < D; ++j) s += Ji[j][i] * dN[a][j];
      gp.g[a][i] = s;
    }
}

// Green-Lagrange strain in Voigt and its first derivative B_a (ns x D per node)
// (dune-localfefunctions greenLagrangeStrains, called at nonlinearelastic.hh:153-155,388-394).
template <int D>
inline void glStrain(const GeoPoint<D>& gp, const double* ue, double F[D][D], double* Ev) {
  constexpr int nn = 1 << D;
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) {
      double h = 0.0;
      for (int a = 0; a < nn; ++a) h += ue[a * D + i] * gp.g[a][j];
      F[i][j] = h + (i == j ? 1.0 : 0.0);
    }
  // E = 1/2 (F^T F - I)
  double E[D][D];
  for (int I = 0; I < D; ++I)
    for (int J = 0; J < D; ++J) {
      double s = 0.0;
      for (int k = 0; k < D; ++k) s += F[k][I] * F[k][J];
      E[I][J] = 0.5 * (s - (I == J ? 1.0 : 0.0));
    }
  if constexpr (D == 3) {
    for (int v = 0; v < 6; ++v) {
      const int i = kVoigt3[v][0], j = kVoigt3[v][1];
      Ev[v] = (v < 3) ? E[i][j] : 2.0 * E[i][j];
    }
  } else {
    Ev[0] = E[0][0];
    Ev[1] = E[1][1];
    Ev[2] = 2.0 * E[0][1];
  }
}

// B_a[v][k] = dEv_v / du_{a,k}: the B-operator of the Green-Lagrange strain in Voigt form, engineering shears
// (nonlinearelastic.hh:313-322 contracts it with the material tangent; strain from Dune::greenLagrangeStrains)
template <int D>
inline void bop(const double F[D][D], const double g[D], double* B /* ns x D */) {
  if constexpr (D == 3) {
    for (int k = 0; k < 3; ++k) {
      B[0 * 3 + k] = F[k][0] * g[0];
      B[1 * 3 + k] = F[k][1] * g[1];
      B[2 * 3 + k] = F[k][2] * g[2];
      B[3 * 3 + k] = F[k][1] * g[2] + F[k][2] * g[1];
      B[4 * 3 + k] = F[k][0] * g[2] + F[k][2] * g[0];
      B[5 * 3 + k] = F[k][0] * g[1] + F[k][1] * g[0];
    }
  } else {
    for (int k = 0; k < 2; ++k) {
      B[0 * 2 + k] = F[k][0] * g[0];
      B[1 * 2 + k] = F[k][1] * g[1];
      B[2 * 2 + k] = F[k][0] * g[1] + F[k][1] * g[0];
    }
  }
}

// Full symmetric PK2 matrix from Voigt stress (no factors on shear: tensorutils.hh:296-326 isStrain=false)
template <int D>
inline void stressMatrix(const double* Sv, double S[D][D]) {
  if constexpr (D == 3) {
    for (int v = 0; v < 6; ++v) {
      const int i = kVoigt3[v][0], j = kVoigt3[v][1];
      S[i][j] = S[j][i] = Sv[v];
    }
  } else {
    S[0][0] = Sv[0];
    S[1][1] = Sv[1];
    S[0][1] = S[1][0] = Sv[2];
  }
}

// ----------------------------------------------------------------------------------------------
// EAS: transformation matrix T and ansatz M (linearandglstrains.hh:22-39,280-328; E4 :68-90;
// helperfunctions.hh:19-25; tensorutils.hh:405-465). Jt = jacobianTransposed, Jt[i][j] = dx_j/dxi_i.
// T[r][c] = (Jt[a][k] Jt[b][l] + Jt[a][l] Jt[b][k]) * (a == b ? 1/2 : 1), (a,b)=pair(r), (k,l)=pair(c).
// ----------------------------------------------------------------------------------------------
template <int D>
inline void easT(const double J[D][D], double* T) {
  constexpr int ns = D * (D + 1) / 2;
  for (int r = 0; r < ns; ++r)
    for (int c = 0; c < ns; ++c) {
      const int a = (D == 3) ? kVoigt3[r][0] : kVoigt2[r][0];
      const int b = (D == 3) ? kVoigt3[r][1] : kVoigt2[r][1];
      const int k = (D == 3) ? kVoigt3[c][0] : kVoigt2[c][0];
      const int l = (D == 3) ? kVoigt3[c][1] : kVoigt2[c][1];
      // Jt[i][j] = J[j][i]
      const double v = J[k][a] * J[l][b] + J[l][a] * J[k][b];
      T[r * ns + c]  = (a == b) ? 0.5 * v : v;
    }
}

// Mtilde (ns x nEAS) in reference coordinates (xi in [0,1]^D); linearandglstrains.hh:68-90 (E4), :280-328 (E21)
template <int D>
inline void easMtilde(int neas, const double xi[D], double* Mt) {
  constexpr int ns = D * (D + 1) / 2;
  for (int i = 0; i < ns * neas; ++i) Mt[i] = 0.0;
  auto at = [&](int r, int c) -> double& { return Mt[r * neas + c]; };
  if constexpr (D == 2) {
    if (neas != 4) throw std::invalid_argument("2D EAS: only 4 modes supported");
    const double x = 2 * xi[0] - 1.0, y = 2 * xi[1] - 1.0;
    at(0, 0) = x;
    at(1, 1) = y;
    at(2, 2) = x;
    at(2, 3) = y;
  } else {
    const double x = 2 * xi[0] - 1.0, y = 2 * xi[1] - 1.0, z = 2 * xi[2] - 1.0;
    if (neas != 9 && neas != 21) throw std::invalid_argument("3D EAS: only 9 or 21 modes supported");
    at(0, 0) = x;
    at(1, 1) = y;
    at(2, 2) = z;
    at(3, 3) = y;
    at(3, 4) = z;
    at(4, 5) = x;
    at(4, 6) = z;
    at(5, 7) = x;
    at(5, 8) = y;
    if (neas == 21) {
      at(3, 9)  = x * y;
      at(3, 10) = x * z;
      at(4, 11) = x * y;
      at(4, 12) = y * z;
      at(5, 13) = x * z;
      at(5, 14) = y * z;
      at(0, 15) = x * y;
      at(0, 16) = x * z;
      at(1, 17) = x * y;
      at(1, 18) = y * z;
      at(2, 19) = x * z;
      at(2, 20) = y * z;
    }
  }
}

// ----------------------------------------------------------------------------------------------
// Problem description
// ----------------------------------------------------------------------------------------------
struct TractionFace {
  int elem, axis, side;  // face of element `elem` normal to reference axis `axis`, at xi_axis = side
};

struct Problem {
  int dim = 3;
  int nn = 0, ne = 0;
  std::vector<double> X;   // nn * dim
  std::vector<int> conn;   // ne * 2^dim
  Material mat;
  int neas   = 0;          // 0 (plain NonLinearElastic), 4 (Q1 E4), 9, 21 (H1)
  int nq1d   = 2;          // Gauss points per direction
  bool faithfulCost = false;  // recompute the tangent per node pair (nonlinearelastic.hh:318) — timing only
  std::vector<uint8_t> fixed;   // ndof
  std::vector<double> ubar;     // ndof, inhomogeneous Dirichlet values per unit load factor (empty if none)
  std::vector<TractionFace> faces;
  double t0[3] = {0, 0, 0};     // traction per unit load factor
  std::vector<std::pair<int, double>> pointLoads;  // R[dof] += coef * lambda
  std::vector<double> alpha;    // ne * neas

  int nodesPerElem() const { return 1 << dim; }
  int64_t ndof() const { return (int64_t)nn * dim; }
};

// ----------------------------------------------------------------------------------------------
// Element residual / tangent (NonLinearElastic + EAS + Traction skills)
// ----------------------------------------------------------------------------------------------
struct ElementOut {
  double K[24 * 24];
  double R[24];
  double energy;
  // EAS intermediates (for the alpha update and for parity tests)
  double Dm[21 * 21], Lm[21 * 24], Rt[21];
};

template <int D>
void elementCompute(const Problem& P, int e, const double* uFull, double lambda, bool wantK, bool wantR,
                    ElementOut& out, const double* alphaOverride = nullptr) {
  constexpr int nn = 1 << D;
  constexpr int nd = nn * D;
  constexpr int ns = D * (D + 1) / 2;
  const int neas   = P.neas;
  double Xe[nn * D], ue[nd];
  for (int a = 0; a < nn; ++a) {
    const int v = P.conn[(size_t)e * nn + a];
    for (int k = 0; k < D; ++k) {
      Xe[a * D + k] = P.X[(size_t)v * D + k];
      ue[a * D + k] = uFull[(size_t)v * D + k];
    }
  }
  const double* alpha = alphaOverride ? alphaOverride : (neas ? &P.alpha[(size_t)e * neas] : nullptr);
  std::fill(out.K, out.K + nd * nd, 0.0);
  std::fill(out.R, out.R + nd, 0.0);
  out.energy = 0.0;
  if (neas) {
    std::fill(out.Dm, out.Dm + neas * neas, 0.0);
    std::fill(out.Lm, out.Lm + neas * nd, 0.0);
    std::fill(out.Rt, out.Rt + neas, 0.0);
  }
  // EAS T0^{-1} at the element centre (helperfunctions.hh:19-25; linearandglstrains.hh:32-34)
  double T0i[36];
  if (neas) {
    double xc[D];
    for (int k = 0; k < D; ++k) xc[k] = 0.5;
    GeoPoint<D> g0;
    geoAt<D>(Xe, xc, g0);
    easT<D>(g0.J, T0i);
    for (int i = 0; i < ns * ns; ++i) T0i[i] *= std::fabs(g0.detJ);
    if (!invertDense(ns, T0i)) throw std::runtime_error("EAS: singular T0");
  }
  const auto rule = tensorGauss<D>(P.nq1d);
  for (const auto& q : rule) {
    GeoPoint<D> gp;
    geoAt<D>(Xe, q.xi, gp);
    const double dV = std::fabs(gp.detJ) * q.w;
    double F[D][D], Ev[ns];
    glStrain<D>(gp, ue, F, Ev);
    double M[6 * 21];
    if (neas) {
      double Mt[6 * 21];
      easMtilde<D>(neas, q.xi, Mt);
      const double s = 1.0 / std::fabs(gp.detJ);
      for (int r = 0; r < ns; ++r)
        for (int c = 0; c < neas; ++c) {
          double acc = 0.0;
          for (int k = 0; k < ns; ++k) acc += T0i[r * ns + k] * Mt[k * neas + c];
          M[r * neas + c] = acc * s;
        }
      for (int r = 0; r < ns; ++r) {
        double acc = 0.0;
        for (int c = 0; c < neas; ++c) acc += M[r * neas + c] * alpha[c];
        Ev[r] += acc;  // enhanced strain E(u) + M alpha (greenlagrangestrain.hh:41-52)
      }
    }
    double Sv[ns], Cv[ns * ns], psi;
    materialD<D>(P.mat, Ev, Sv, Cv, &psi);
    out.energy += psi * dV;
    double B[nn][ns * D];
    for (int a = 0; a < nn; ++a) bop<D>(F, gp.g[a], B[a]);
    if (wantR) {
      for (int a = 0; a < nn; ++a)
        for (int k = 0; k < D; ++k) {
          double s = 0.0;
          for (int v = 0; v < ns; ++v) s += B[a][v * D + k] * Sv[v];
          out.R[a * D + k] += s * dV;  // internalForcesFunction (nonlinearelastic.hh:335-342)
        }
    }
    if (wantK || neas) {
      double Sm[D][D];
      stressMatrix<D>(Sv, Sm);
      if (wantK) {
        // C*B_b once per Gauss point (or per node pair in the reference-cost mode, nonlinearelastic.hh:318)
        double CB[nn][ns * D];
        for (int b = 0; b < nn; ++b)
          for (int v = 0; v < ns; ++v)
            for (int l = 0; l < D; ++l) {
              double cb = 0.0;
              for (int w = 0; w < ns; ++w) cb += Cv[v * ns + w] * B[b][w * D + l];
              CB[b][v * D + l] = cb;
            }
        for (int a = 0; a < nn; ++a) {
          for (int b = 0; b < nn; ++b) {
            const double* CBb = CB[b];
            double CBloc[ns * D];
            if (P.faithfulCost) {
              double Cloc[ns * ns], Stmp[ns];
              materialD<D>(P.mat, Ev, Stmp, Cloc, nullptr);
              for (int v = 0; v < ns; ++v)
                for (int l = 0; l < D; ++l) {
                  double cb = 0.0;
                  for (int w = 0; w < ns; ++w) cb += Cloc[v * ns + w] * B[b][w * D + l];
                  CBloc[v * D + l] = cb;
                }
              CBb = CBloc;
            }
            // material part B_a^T C B_b (nonlinearelastic.hh:313-322)
            for (int k = 0; k < D; ++k)
              for (int l = 0; l < D; ++l) {
                double s = 0.0;
                for (int v = 0; v < ns; ++v) s += B[a][v * D + k] * CBb[v * D + l];
                out.K[(a * D + k) * nd + (b * D + l)] += s * dV;
              }
            // geometric part g_a^T S g_b delta_kl (nonlinearelastic.hh:294-302)
            double kg = 0.0;
            for (int i = 0; i < D; ++i)
              for (int j = 0; j < D; ++j) kg += gp.g[a][i] * Sm[i][j] * gp.g[b][j];
            for (int k = 0; k < D; ++k) out.K[(a * D + k) * nd + (b * D + k)] += kg * dV;
          }
        }
      }
      if (neas) {
        // D += M^T C M dV ; L += M^T C B dV ; Rt += M^T S dV (enhancedassumedstrains.hh:378-434)
        double CM[6 * 21];
        for (int v = 0; v < ns; ++v)
          for (int c = 0; c < neas; ++c) {
            double s = 0.0;
            for (int w = 0; w < ns; ++w) s += Cv[v * ns + w] * M[w * neas + c];
            CM[v * neas + c] = s;
          }
        for (int i = 0; i < neas; ++i) {
          for (int j = 0; j < neas; ++j) {
            double s = 0.0;
            for (int v = 0; v < ns; ++v) s += M[v * neas + i] * CM[v * neas + j];
            out.Dm[i * neas + j] += s * dV;
          }
          for (int a = 0; a < nn; ++a)
            for (int k = 0; k < D; ++k) {
              double s = 0.0;
              for (int v = 0; v < ns; ++v) s += CM[v * neas + i] * B[a][v * D + k];
              out.Lm[i * nd + a * D + k] += s * dV;
            }
          double s = 0.0;
          for (int v = 0; v < ns; ++v) s += M[v * neas + i] * Sv[v];
          out.Rt[i] += s * dV;
        }
      }
    }
  }
  if (neas) {
    // Condensation: K -= L^T D^{-1} L (upper), lower := upper^T; R -= L^T D^{-1} Rt
    // (enhancedassumedstrains.hh:286-296, 336-345)
    std::vector<double> Di(out.Dm, out.Dm + neas * neas);
    if (!invertDense(neas, Di.data())) throw std::runtime_error("EAS: singular D");
    std::vector<double> DiL(neas * nd), DiR(neas);
    for (int i = 0; i < neas; ++i) {
      for (int c = 0; c < nd; ++c) {
        double s = 0.0;
        for (int j = 0; j < neas; ++j) s += Di[i * neas + j] * out.Lm[j * nd + c];
        DiL[i * nd + c] = s;
      }
      double s = 0.0;
      for (int j = 0; j < neas; ++j) s += Di[i * neas + j] * out.Rt[j];
      DiR[i] = s;
    }
    if (wantK) {
      for (int r = 0; r < nd; ++r)
        for (int c = r; c < nd; ++c) {
          double s = 0.0;
          for (int i = 0; i < neas; ++i) s += out.Lm[i * nd + r] * DiL[i * nd + c];
          out.K[r * nd + c] -= s;
        }
      for (int r = 0; r < nd; ++r)
        for (int c = 0; c < r; ++c) out.K[r * nd + c] = out.K[c * nd + r];
    }
    if (wantR)
      for (int r = 0; r < nd; ++r) {
        double s = 0.0;
        for (int i = 0; i < neas; ++i) s += out.Lm[i * nd + r] * DiR[i];
        out.R[r] -= s;
      }
  }
  // Traction (traction.hh:71-138): 1-point face rule, R_a -= N_a t w |J_Gamma|, energy -= t.u w |J_Gamma|
  if (!P.faces.empty()) {
    for (const auto& f : P.faces) {
      if (f.elem != e) continue;
      double xi[D];
      for (int k = 0; k < D; ++k) xi[k] = 0.5;
      xi[f.axis] = (double)f.side;
      GeoPoint<D> gp;
      geoAt<D>(Xe, xi, gp);
      // face measure: |J restricted to the in-face reference directions|
      double meas;
      if constexpr (D == 3) {
        int t[2], n = 0;
        for (int k = 0; k < 3; ++k)
          if (k != f.axis) t[n++] = k;
        const double a0[3] = {gp.J[0][t[0]], gp.J[1][t[0]], gp.J[2][t[0]]};
        const double a1[3] = {gp.J[0][t[1]], gp.J[1][t[1]], gp.J[2][t[1]]};
        const double cx = a0[1] * a1[2] - a0[2] * a1[1], cy = a0[2] * a1[0] - a0[0] * a1[2],
                     cz = a0[0] * a1[1] - a0[1] * a1[0];
        meas = std::sqrt(cx * cx + cy * cy + cz * cz);
      } else {
        const int t = 1 - f.axis;
        meas = std::sqrt(gp.J[0][t] * gp.J[0][t] + gp.J[1][t] * gp.J[1][t]);
      }
      double tu = 0.0;
      for (int a = 0; a < nn; ++a)
        for (int k = 0; k < D; ++k) tu += gp.N[a] * ue[a * D + k] * (P.t0[k] * lambda);
      out.energy -= tu * meas;
      if (wantR)
        for (int a = 0; a < nn; ++a)
          for (int k = 0; k < D; ++k) out.R[a * D + k] -= gp.N[a] * (P.t0[k] * lambda) * meas;
    }
  }
}

// Stress at the local point xi of a 3D plain element (NonLinearElastic::calculateAtImpl,
// nonlinearelastic.hh:237-270): kind 0 PK2 S, 1 Kirchhoff F S F^T, 2 Cauchy F S F^T / det F; Voigt
// [11,22,33,23,13,12] without shear factors (toVoigt(., false))
inline void stressAt3(const Problem& P, int e, const double* uFull, const double xi[3], int kind, double out[6]) {
  double Xe[24], ue[24];
  for (int a = 0; a < 8; ++a) {
    const int v = P.conn[(size_t)e * 8 + a];
    for (int k = 0; k < 3; ++k) {
      Xe[a * 3 + k] = P.X[(size_t)v * 3 + k];
      ue[a * 3 + k] = uFull[(size_t)v * 3 + k];
    }
  }
  GeoPoint<3> gp;
  geoAt<3>(Xe, xi, gp);
  double F[3][3], Ev[6], S[6];
  glStrain<3>(gp, ue, F, Ev);
  material3d(P.mat, Ev, S, nullptr, nullptr);
  if (kind == 0) {
    for (int v = 0; v < 6; ++v) out[v] = S[v];
    return;
  }
  double Sm[3][3], T[3][3];
  for (int v = 0; v < 6; ++v) Sm[kVoigt3[v][0]][kVoigt3[v][1]] = Sm[kVoigt3[v][1]][kVoigt3[v][0]] = S[v];
  const double detF = F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) -
                      F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
                      F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double t = 0.0;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) t += F[i][k] * Sm[k][l] * F[j][l];
      T[i][j] = kind == 2 ? t / detF : t;
    }
  for (int v = 0; v < 6; ++v) out[v] = T[kVoigt3[v][0]][kVoigt3[v][1]];
}

inline void elementComputeDispatch(const Problem& P, int e, const double* u, double lambda, bool wantK, bool wantR,
                                   ElementOut& out, const double* alphaOverride = nullptr) {
  if (P.dim == 3)
    elementCompute<3>(P, e, u, lambda, wantK, wantR, out, alphaOverride);
  else
    elementCompute<2>(P, e, u, lambda, wantK, wantR, out, alphaOverride);
}

// ----------------------------------------------------------------------------------------------
// Sparsity pattern (Eigen ColMajor CSC, sorted inner indices) and element slot maps
// (simpleassemblers.inl:207-222, 225-251, 254-287; eigensparseaddon.hh:19-29)
// ----------------------------------------------------------------------------------------------
struct CSC {
  int64_t n = 0;
  std::vector<int32_t> outer;  // n + 1
  std::vector<int32_t> inner;  // nnz
  std::vector<std::vector<int64_t>> slots;  // per element: (nd)^2 (full) in column-major local order
};

inline std::vector<int64_t> elementDofs(const Problem& P, int e) {
  const int nn = P.nodesPerElem();
  std::vector<int64_t> d;
  for (int a = 0; a < nn; ++a)
    for (int k = 0; k < P.dim; ++k) d.push_back((int64_t)P.conn[(size_t)e * nn + a] * P.dim + k);
  return d;
}

inline std::vector<int64_t> constraintsBelow(const Problem& P) {
  std::vector<int64_t> cb(P.ndof());
  int64_t c = 0;
  for (int64_t i = 0; i < P.ndof(); ++i) {
    cb[i] = c;
    if (P.fixed[i]) ++c;
  }
  return cb;
}

inline CSC buildPattern(const Problem& P, bool reduced) {
  const int64_t ndof = P.ndof();
  std::vector<int64_t> cb;
  int64_t n = ndof;
  if (reduced) {
    cb = constraintsBelow(P);
    n  = ndof - std::count(P.fixed.begin(), P.fixed.end(), (uint8_t)1);
  }
  // triplets (row, col)
  std::vector<std::pair<int64_t, int64_t>> trip;  // (col, row) for column-major sort
  for (int e = 0; e < P.ne; ++e) {
    const auto d = elementDofs(P, e);
    for (auto r : d)
      for (auto c : d) {
        if (reduced && (P.fixed[r] || P.fixed[c])) continue;
        const int64_t rr = reduced ? r - cb[r] : r, cc = reduced ? c - cb[c] : c;
        trip.emplace_back(cc, rr);
      }
  }
  std::sort(trip.begin(), trip.end());
  trip.erase(std::unique(trip.begin(), trip.end()), trip.end());
  CSC m;
  m.n = n;
  m.outer.assign(n + 1, 0);
  m.inner.resize(trip.size());
  for (size_t i = 0; i < trip.size(); ++i) {
    m.outer[trip[i].first + 1]++;
    m.inner[i] = (int32_t)trip[i].second;
  }
  for (int64_t c = 0; c < n; ++c) m.outer[c + 1] += m.outer[c];
  auto linIdx = [&](int64_t r, int64_t c) -> int64_t {
    const auto b = m.inner.begin() + m.outer[c], en = m.inner.begin() + m.outer[c + 1];
    const auto it = std::lower_bound(b, en, (int32_t)r);
    return it - m.inner.begin();
  };
  m.slots.resize(P.ne);
  for (int e = 0; e < P.ne; ++e) {
    const auto d = elementDofs(P, e);
    auto& s = m.slots[e];
    if (!reduced) {
      // slot[c*n + r] (column-major over the local pair)
      for (auto c : d)
        for (auto r : d) s.push_back(linIdx(r, c));
    } else {
      // row-major over unconstrained local pairs (simpleassemblers.inl:268-287)
      for (auto r : d) {
        if (P.fixed[r]) continue;
        for (auto c : d) {
          if (P.fixed[c]) continue;
          s.push_back(linIdx(r - cb[r], c - cb[c]));
        }
      }
    }
  }
  return m;
}

// ----------------------------------------------------------------------------------------------
// Assembly (simpleassemblers.inl:60-169)
// ----------------------------------------------------------------------------------------------
enum DBCOption { Raw = 0, Full = 1, Reduced = 2 };

struct Assembler {
  Problem* P = nullptr;
  CSC pat, patRed;
  bool havePat = false, havePatRed = false;
  std::vector<double> vals, valsRed;
  std::vector<int64_t> cb;

  explicit Assembler(Problem* p) : P(p) { cb = constraintsBelow(*p); }

  // Optimized CPU timing variant (not the reference's cost structure): with threads > 1 the element loops run
  // color by color over a greedy coloring in which no two elements of a color share a node, one OpenMP loop per
  // color, so the scatter needs no atomics. Same sums per entry as the serial loop, in another order.
  int threads = 1;
  std::vector<std::vector<int>> colors;
  void ensureColors() {
    if (!colors.empty()) return;
    const int npe = P->nodesPerElem();
    std::vector<uint64_t> used(P->nn, 0);
    for (int e = 0; e < P->ne; ++e) {
      uint64_t m = 0;
      for (int a = 0; a < npe; ++a) m |= used[P->conn[(size_t)e * npe + a]];
      int c = 0;
      while (c < 64 && ((m >> c) & 1)) ++c;
      if (c == 64) throw std::runtime_error("element coloring needs more than 64 colors");
      if ((int)colors.size() <= c) colors.resize(c + 1);
      colors[c].push_back(e);
      for (int a = 0; a < npe; ++a) used[P->conn[(size_t)e * npe + a]] |= 1ull << c;
    }
  }
  // f(e, eo) for every element: serial in element order, or color by color on `threads` threads
  template <class F>
  void forElements(bool wantK, bool wantR, const double* u, double lambda, F&& f) {
    if (threads <= 1) {
      ElementOut eo;
      for (int e = 0; e < P->ne; ++e) {
        elementComputeDispatch(*P, e, u, lambda, wantK, wantR, eo);
        f(e, eo);
      }
      return;
    }
    ensureColors();
    std::exception_ptr err;  // an exception may not leave an OpenMP region: the first one is rethrown after it
    for (const auto& col : colors) {
      const int nc = (int)col.size();
#pragma omp parallel num_threads(threads)
      {
        ElementOut eo;
#pragma omp for schedule(static)
        for (int i = 0; i < nc; ++i) {
          try {
            elementComputeDispatch(*P, col[i], u, lambda, wantK, wantR, eo);
            f(col[i], eo);
          } catch (...) {
#pragma omp critical
            if (!err) err = std::current_exception();
          }
        }
      }
      if (err) std::rethrow_exception(err);
    }
  }

  const CSC& pattern(bool reduced) {
    if (reduced) {
      if (!havePatRed) {
        patRed     = buildPattern(*P, true);
        havePatRed = true;
      }
      return patRed;
    }
    if (!havePat) {
      pat     = buildPattern(*P, false);
      havePat = true;
    }
    return pat;
  }

  // K in CSC, returns values vector sized nnz of the corresponding pattern
  std::vector<double>& matrix(const double* u, double lambda, int dbc) {
    const int nd = P->nodesPerElem() * P->dim;
    ElementOut eo;
    if (dbc == Reduced) {
      const CSC& m = pattern(true);
      valsRed.assign(m.inner.size(), 0.0);
      for (int e = 0; e < P->ne; ++e) {
        elementComputeDispatch(*P, e, u, lambda, true, false, eo);
        const auto d = elementDofs(*P, e);
        size_t li    = 0;
        for (int r = 0; r < nd; ++r) {
          if (P->fixed[d[r]]) continue;
          for (int c = 0; c < nd; ++c) {
            if (P->fixed[d[c]]) continue;
            valsRed[m.slots[e][li++]] += eo.K[r * nd + c];
          }
        }
      }
      return valsRed;
    }
    const CSC& m = pattern(false);
    vals.assign(m.inner.size(), 0.0);
    forElements(true, false, u, lambda, [&](int e, const ElementOut& o) {
      size_t li = 0;
      for (int c = 0; c < nd; ++c)
        for (int r = 0; r < nd; ++r) vals[m.slots[e][li++]] += o.K[r * nd + c];
    });
    if (dbc == Full) {
      for (int64_t c = 0; c < m.n; ++c)
        for (int32_t p = m.outer[c]; p < m.outer[c + 1]; ++p) {
          const int64_t r = m.inner[p];
          if (P->fixed[r] || P->fixed[c]) vals[p] = (r == c && P->fixed[r]) ? 1.0 : 0.0;
        }
    }
    return vals;
  }

  std::vector<double> vector(const double* u, double lambda, int dbc) {
    const int nd = P->nodesPerElem() * P->dim;
    std::vector<double> R(P->ndof(), 0.0);
    forElements(false, true, u, lambda, [&](int e, const ElementOut& o) {
      const auto d = elementDofs(*P, e);
      for (int i = 0; i < nd; ++i) R[d[i]] += o.R[i];
    });
    // Point loads are AssemblerManipulator callbacks: they run after the DBC treatment of the wrapped
    // assembler (assemblermanipulatorbuildingblocks.hh:103-122); reduced vectors take the reduced index.
    if (dbc == Raw || dbc == Full) {
      if (dbc == Full)
        for (int64_t i = 0; i < P->ndof(); ++i)
          if (P->fixed[i]) R[i] = 0.0;
      for (auto& pl : P->pointLoads) R[pl.first] += pl.second * lambda;
      return R;
    }
    std::vector<double> Rr;
    for (int64_t i = 0; i < P->ndof(); ++i)
      if (!P->fixed[i]) Rr.push_back(R[i]);
    for (auto& pl : P->pointLoads)
      if (!P->fixed[pl.first]) Rr[pl.first - cb[pl.first]] += pl.second * lambda;
    return Rr;
  }

  double energy(const double* u, double lambda) {
    ElementOut eo;
    double s = 0.0;
    for (int e = 0; e < P->ne; ++e) {
      elementComputeDispatch(*P, e, u, lambda, false, false, eo);
      s += eo.energy;
    }
    return s;
  }
};

// ----------------------------------------------------------------------------------------------
// Linear solvers
// ----------------------------------------------------------------------------------------------
// CSC (symmetric storage of the full matrix) SpMV y = A x
inline void cscMatVec(const CSC& m, const std::vector<double>& v, const double* x, double* y) {
  std::fill(y, y + m.n, 0.0);
  for (int64_t c = 0; c < m.n; ++c) {
    const double xc = x[c];
    for (int32_t p = m.outer[c]; p < m.outer[c + 1]; ++p) y[m.inner[p]] += v[p] * xc;
  }
}

// Banded LDL^T direct solve (stand-in for the reference's sparse direct solvers: sd_SimplicialLDLT,
// sd_UmfPackLU — linearsolver.cpp:35-36,47-48). Exact up to rounding for SPD K.
struct BandedLDLT {
  int64_t n = 0, bw = 0;
  std::vector<double> a;  // row i: entries j in [i-bw, i], stored a[i*(bw+1) + (j - i + bw)]
  bool factorize(const CSC& m, const std::vector<double>& v) {
    n  = m.n;
    bw = 0;
    for (int64_t c = 0; c < n; ++c)
      for (int32_t p = m.outer[c]; p < m.outer[c + 1]; ++p) bw = std::max<int64_t>(bw, std::abs(m.inner[p] - c));
    a.assign((size_t)n * (bw + 1), 0.0);
    for (int64_t c = 0; c < n; ++c)
      for (int32_t p = m.outer[c]; p < m.outer[c + 1]; ++p) {
        const int64_t r = m.inner[p];
        if (c <= r) a[(size_t)r * (bw + 1) + (c - r + bw)] = v[p];
      }
    auto A = [&](int64_t i, int64_t j) -> double& { return a[(size_t)i * (bw + 1) + (j - i + bw)]; };
    for (int64_t j = 0; j < n; ++j) {
      // d_j = a_jj - sum_k L_jk^2 d_k
      const int64_t k0 = std::max<int64_t>(0, j - bw);
      double d         = A(j, j);
      for (int64_t k = k0; k < j; ++k) d -= A(j, k) * A(j, k) * A(k, k);
      if (d == 0.0 || !std::isfinite(d)) return false;
      A(j, j) = d;
      for (int64_t i = j + 1; i <= std::min(n - 1, j + bw); ++i) {
        double s         = A(i, j);
        const int64_t kk = std::max<int64_t>(0, i - bw);
        for (int64_t k = std::max(k0, kk); k < j; ++k) s -= A(i, k) * A(j, k) * A(k, k);
        A(i, j) = s / d;
      }
    }
    return true;
  }
  void solve(const double* b, double* x) const {
    auto A = [&](int64_t i, int64_t j) -> double { return a[(size_t)i * (bw + 1) + (j - i + bw)]; };
    for (int64_t i = 0; i < n; ++i) {
      double s = b[i];
      for (int64_t k = std::max<int64_t>(0, i - bw); k < i; ++k) s -= A(i, k) * x[k];
      x[i] = s;
    }
    for (int64_t i = 0; i < n; ++i) x[i] /= A(i, i);
    for (int64_t i = n - 1; i >= 0; --i) {
      double s = x[i];
      for (int64_t k = i + 1; k <= std::min(n - 1, i + bw); ++k) s -= A(k, i) * x[k];
      x[i] = s;
    }
  }
};

// Jacobi-preconditioned CG with Eigen ConjugateGradient semantics (linearsolver.cpp:23-24):
// x0 = 0, stop when ||r||^2 < tol^2 ||b||^2, at most maxIter iterations (Eigen default 2n).
struct CGInfo {
  int iterations = 0;
  double error   = 0.0;
};
inline CGInfo jacobiCG(const CSC& m, const std::vector<double>& v, const double* b, double* x, double tol,
                       int maxIter, int nthreads = 1) {
  const int64_t n = m.n;
  std::vector<double> diag(n, 0.0), r(n), z(n), p(n), q(n);
  for (int64_t c = 0; c < n; ++c)
    for (int32_t k = m.outer[c]; k < m.outer[c + 1]; ++k)
      if (m.inner[k] == c) diag[c] = v[k];
  for (int64_t i = 0; i < n; ++i) diag[i] = diag[i] != 0.0 ? 1.0 / diag[i] : 1.0;
  // y = A x for the symmetric matrix: use column access == row access
  auto matvec = [&](const double* xx, double* yy) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t c = 0; c < n; ++c) {  // row c of a symmetric matrix = column c
      double s = 0.0;
      for (int32_t k = m.outer[c]; k < m.outer[c + 1]; ++k) s += v[k] * xx[m.inner[k]];
      yy[c] = s;
    }
  };
  double rhsNorm2 = 0.0;
  for (int64_t i = 0; i < n; ++i) rhsNorm2 += b[i] * b[i];
  CGInfo info;
  std::fill(x, x + n, 0.0);
  if (rhsNorm2 == 0.0) return info;
  const double threshold = std::max(tol * tol * rhsNorm2, std::numeric_limits<double>::min());
  for (int64_t i = 0; i < n; ++i) r[i] = b[i];
  double rn2 = rhsNorm2;
  if (rn2 < threshold) {
    info.error = std::sqrt(rn2 / rhsNorm2);
    return info;
  }
  for (int64_t i = 0; i < n; ++i) p[i] = diag[i] * r[i];
  double absNew = 0.0;
  for (int64_t i = 0; i < n; ++i) absNew += r[i] * p[i];
  int it = 0;
  while (it < maxIter) {
    matvec(p.data(), q.data());
    double pq = 0.0;
#pragma omp parallel for num_threads(nthreads) reduction(+ : pq)
    for (int64_t i = 0; i < n; ++i) pq += p[i] * q[i];
    const double alpha = absNew / pq;
    rn2                = 0.0;
#pragma omp parallel for num_threads(nthreads) reduction(+ : rn2)
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      rn2 += r[i] * r[i];
    }
    if (rn2 < threshold) break;
    double absOld = absNew;
    absNew        = 0.0;
#pragma omp parallel for num_threads(nthreads) reduction(+ : absNew)
    for (int64_t i = 0; i < n; ++i) {
      z[i] = diag[i] * r[i];
      absNew += r[i] * z[i];
    }
    const double beta = absNew / absOld;
#pragma omp parallel for num_threads(nthreads)
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    ++it;
  }
  info.iterations = it;
  info.error      = std::sqrt(rn2 / rhsNorm2);
  return info;
}

// ----------------------------------------------------------------------------------------------
// Newton-Raphson and LoadControl (newtonraphson.hh:196-257; loadcontrol.inl:21-57;
// nonlinearsolverfactory.hh:33-56; functionhelper.hh:170-184)
// ----------------------------------------------------------------------------------------------
enum LinSolverKind { DirectLDLT = 0, JacobiCGSolver = 1 };
enum PredictorKind { PredictConstant = 0, PredictLinear = 1 };

struct NRSettings {
  double tol  = 1e-8;
  int maxIter = 20;
  int minIter = 0;
  int linSolver = DirectLDLT;
  double cgTol  = 2.220446049250313e-16;  // Eigen default: NumTraits<double>::epsilon()
  int cgMaxIter = -1;                     // Eigen default: 2n
  int threads   = 1;
};

struct NRInfo {
  bool success = false;
  int iterations = 0;
  double residualNorm = 0, correctionNorm = 0;
  long cgIterations = 0;
};

inline double norm2(const std::vector<double>& v) {
  double s = 0.0;
  for (double x : v) s += x * x;
  return std::sqrt(s);
}

struct Solver {
  Problem* P;
  Assembler asmb;
  NRSettings nr;
  explicit Solver(Problem* p) : P(p), asmb(p) {}

  void linearSolve(const std::vector<double>& Kv, const std::vector<double>& rhs, std::vector<double>& x,
                   NRInfo& info) {
    const CSC& m = asmb.pattern(false);
    x.assign(m.n, 0.0);
    if (nr.linSolver == DirectLDLT) {
      BandedLDLT f;
      if (!f.factorize(m, Kv)) throw std::runtime_error("LDLT factorization failed");
      f.solve(rhs.data(), x.data());
    } else {
      const int maxIt = nr.cgMaxIter > 0 ? nr.cgMaxIter : (int)(2 * m.n);
      auto ci         = jacobiCG(m, Kv, rhs.data(), x.data(), nr.cgTol, maxIt, nr.threads);
      info.cgIterations += ci.iterations;
    }
  }

  void alphaUpdate(const std::vector<double>& u, double lambda, const std::vector<double>& du) {
    if (!P->neas) return;
    // alpha -= D^{-1} (Rt + L du_e) with D, L, Rt at the pre-update state (enhancedassumedstrains.hh:225-248)
    const int nd = P->nodesPerElem() * P->dim, ne = P->neas;
    ElementOut eo;
    std::vector<double> newAlpha = P->alpha;
    for (int e = 0; e < P->ne; ++e) {
      elementComputeDispatch(*P, e, u.data(), lambda, false, false, eo);
      const auto d = elementDofs(*P, e);
      std::vector<double> rhs(ne), Di(eo.Dm, eo.Dm + ne * ne);
      for (int i = 0; i < ne; ++i) {
        double s = eo.Rt[i];
        for (int c = 0; c < nd; ++c) s += eo.Lm[i * nd + c] * du[d[c]];
        rhs[i] = s;
      }
      if (!invertDense(ne, Di.data())) throw std::runtime_error("EAS: singular D in update");
      for (int i = 0; i < ne; ++i) {
        double s = 0.0;
        for (int j = 0; j < ne; ++j) s += Di[i * ne + j] * rhs[j];
        newAlpha[(size_t)e * ne + i] -= s;
      }
    }
    P->alpha.swap(newAlpha);
  }

  // IDBC force: K_raw(x) * ubar, zeroed at constrained dofs (functionhelper.hh:170-184)
  std::vector<double> idbcForce(const std::vector<double>& u, double lambda) {
    const CSC& m   = asmb.pattern(false);
    auto& Kraw     = asmb.matrix(u.data(), lambda, Raw);
    std::vector<double> f(m.n);
    cscMatVec(m, Kraw, P->ubar.data(), f.data());
    for (int64_t i = 0; i < m.n; ++i)
      if (P->fixed[i]) f[i] = 0.0;
    return f;
  }

  void syncIDBC(std::vector<double>& u, double lambda) {
    if (P->ubar.empty()) return;
    for (int64_t i = 0; i < P->ndof(); ++i) {
      const double v = P->ubar[i] * lambda;
      if (v != 0.0) u[i] = v;
    }
  }

  NRInfo solve(std::vector<double>& u, double& lambda, double stepSize) {
    NRInfo info;
    info.success   = true;
    auto rx        = asmb.vector(u.data(), lambda, Full);
    auto Ax        = asmb.matrix(u.data(), lambda, Full);
    double rNorm   = norm2(rx);
    if (!P->ubar.empty()) {
      auto f = idbcForce(u, lambda);
      for (size_t i = 0; i < rx.size(); ++i) rx[i] += f[i] * stepSize;
      rNorm = norm2(rx);
    }
    info.residualNorm = rNorm;
    int iter          = 0;
    std::vector<double> corr, mrx(rx.size());
    while ((rNorm > nr.tol && iter < nr.maxIter) || iter < nr.minIter) {
      for (size_t i = 0; i < rx.size(); ++i) mrx[i] = -rx[i];
      linearSolve(Ax, mrx, corr, info);
      info.correctionNorm = norm2(corr);
      alphaUpdate(u, lambda, corr);  // CORRECTION_UPDATED (newtonraphson.hh:230)
      for (size_t i = 0; i < u.size(); ++i) u[i] += corr[i];
      syncIDBC(u, lambda);
      rx                = asmb.vector(u.data(), lambda, Full);
      Ax                = asmb.matrix(u.data(), lambda, Full);
      rNorm             = norm2(rx);
      info.residualNorm = rNorm;
      ++iter;
    }
    if (iter == nr.maxIter) info.success = false;
    info.iterations = iter;
    return info;
  }
};

struct ControlInfo {
  bool success = false;
  int totalIterations = 0;
  std::vector<int> iterations;  // per solve, including the initial one
  std::vector<double> residualNorms;
  long cgIterations = 0;
};

// LoadControl::run; predictor = constant (reference, loadcontrol.inl:55-57) or linear extrapolation
// u += (u_n - u_{n-1}) before lambda += dlambda (ArcLength-style secant, pathfollowingfunctions.hh:188-195).
inline ControlInfo loadControl(Solver& S, std::vector<double>& u, double& lambda, int steps, double tbegin,
                               double tend, int predictor = PredictConstant) {
  ControlInfo ci;
  const double dl = (tend - tbegin) / steps;
  auto info       = S.solve(u, lambda, 0.0);
  ci.iterations.push_back(info.iterations);
  ci.residualNorms.push_back(info.residualNorm);
  ci.totalIterations += info.iterations;
  ci.cgIterations += info.cgIterations;
  if (!info.success) return ci;
  std::vector<double> uPrev = u;
  for (int ls = 0; ls < steps; ++ls) {
    if (predictor == PredictLinear && ls > 0) {
      for (size_t i = 0; i < u.size(); ++i) {
        const double inc = u[i] - uPrev[i];
        uPrev[i]         = u[i];
        u[i] += inc;
      }
    } else {
      uPrev = u;
    }
    lambda += dl;
    info = S.solve(u, lambda, dl);
    if (info.iterations == 0) S.syncIDBC(u, lambda);
    ci.iterations.push_back(info.iterations);
    ci.residualNorms.push_back(info.residualNorm);
    ci.totalIterations += info.iterations;
    ci.cgIterations += info.cgIterations;
    if (!info.success) return ci;
  }
  ci.success = true;
  return ci;
}

// ----------------------------------------------------------------------------------------------
// Structured meshes (YaspGrid ordering: vertex (i,j,k) -> i + (nx+1)(j + (ny+1)k), element likewise)
// with seeded interior jitter (SplitMix64 -> 53-bit uniform double).
// ----------------------------------------------------------------------------------------------
struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z          = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

inline void structuredMesh(int dim, const int* nel, const double* L, double jitter, uint64_t seed, Problem& P) {
  P.dim = dim;
  int nv[3] = {1, 1, 1}, ne[3] = {1, 1, 1};
  for (int k = 0; k < dim; ++k) {
    nv[k] = nel[k] + 1;
    ne[k] = nel[k];
  }
  P.nn = nv[0] * nv[1] * nv[2];
  P.ne = ne[0] * ne[1] * ne[2];
  P.X.assign((size_t)P.nn * dim, 0.0);
  SplitMix64 rng(seed);
  for (int v = 0; v < P.nn; ++v) {
    const int idx[3] = {v % nv[0], (v / nv[0]) % nv[1], v / (nv[0] * nv[1])};
    bool interior    = true;
    for (int k = 0; k < dim; ++k)
      if (idx[k] == 0 || idx[k] == nel[k]) interior = false;
    for (int k = 0; k < dim; ++k) {
      const double h = L[k] / nel[k];
      double x       = idx[k] * h;
      if (jitter > 0.0 && interior) x += (2.0 * rng.uniform() - 1.0) * jitter * h;
      P.X[(size_t)v * dim + k] = x;
    }
  }
  const int nc = 1 << dim;
  P.conn.assign((size_t)P.ne * nc, 0);
  for (int e = 0; e < P.ne; ++e) {
    const int i = e % ne[0], j = (e / ne[0]) % ne[1], k = e / (ne[0] * ne[1]);
    for (int a = 0; a < nc; ++a) {
      const int di = a & 1, dj = (a >> 1) & 1, dk = (a >> 2) & 1;
      P.conn[(size_t)e * nc + a] = (i + di) + nv[0] * ((j + dj) + nv[1] * (k + dk));
    }
  }
  P.fixed.assign((size_t)P.nn * dim, 0);
}

}  // namespace ikoracle
