// eas.hip — EAS (enhanced Green-Lagrange strain, 21 modes) hex8 element: residual, condensed tangent and
// the internal-variable update.
//
// Restates EnhancedAssumedStrains (ikarus/finiteelements/mechanics/enhancedassumedstrains.hh:225-434)
// with the E21 ansatz (strainenhancements/easvariants/linearandglstrains.hh:22-39,280-328),
// T0 = T(xi_0) * detJ(xi_0) (easvariants/helperfunctions.hh:19-25; T from ikarus/utils/tensorutils.hh:440-465)
// and the GreenLagrangeStrain enhancement (easfunctions/greenlagrangestrain.hh:41-141):
//   E_enh = E(u) + M alpha,  M(xi) = T0^{-1} Mtilde(xi) / detJ(xi)
//   K_uu = sum (B^T C B + kg) dV,  R_u = sum B^T S dV,  D = sum M^T C M dV,  L = sum M^T C B dV,
//   Rt = sum M^T S dV,  K = K_uu - L^T D^{-1} L (upper, mirrored),  R = R_u - L^T D^{-1} Rt,
//   alpha <- alpha - D^{-1}(Rt + L du_e) on CORRECTION_UPDATED, with D, L, Rt of the state the tangent
//   was assembled at (the reference recomputes them at that same state).
//
// One 64-lane workgroup per element. Lanes 0-7 evaluate the 8 Gauss points (kinematics, enhanced strain,
// S, C, M) into LDS; all 64 lanes then accumulate the 36 upper 3x3 blocks of K_uu, D, L, Rt and R_u per
// Gauss point from LDS; the condensation is a right-looking LDL^T of D in LDS followed by 25
// forward/backward substitutions (columns of L and Rt) and the rank-21 update of K and R. Outputs go to
// the same element-by-element staging as the plain element (36 upper blocks + 24 residual entries), plus
// the cached Y = D^{-1} L (21x24) and y = D^{-1} Rt (21) used by the alpha update.
#include "ik_internal.hh"

namespace ik {
namespace {

constexpr int NE = 21;  // enhanced modes

__host__ __device__ constexpr int vI(int v) { return v < 3 ? v : (v == 3 ? 1 : 0); }
__host__ __device__ constexpr int vJ(int v) { return v < 3 ? v : (v == 5 ? 1 : 2); }
__host__ __device__ constexpr int sym6(int p, int q) {
  return p <= q ? p * 6 - (p * (p - 1)) / 2 + (q - p) : q * 6 - (q * (q - 1)) / 2 + (p - q);
}
__host__ __device__ constexpr int upperBlock(int a, int b) { return 8 * a - (a * (a - 1)) / 2 + (b - a); }

// E21 ansatz in reference coordinates: column c has one nonzero, in Voigt row kRow[c], of value
// f_c(x, y, z) with x = 2 xi - 1 etc. (linearandglstrains.hh:300-322)
__constant__ int kRow[NE] = {0, 1, 2, 3, 3, 4, 4, 5, 5, 3, 3, 4, 4, 5, 5, 0, 0, 1, 1, 2, 2};
__device__ inline double ansatz(int c, double x, double y, double z) {
  switch (c) {
    case 0: return x;
    case 1: return y;
    case 2: return z;
    case 3: return y;
    case 4: return z;
    case 5: return x;
    case 6: return z;
    case 7: return x;
    case 8: return y;
    case 9: return x * y;
    case 10: return x * z;
    case 11: return x * y;
    case 12: return y * z;
    case 13: return x * z;
    case 14: return y * z;
    case 15: return x * y;
    case 16: return x * z;
    case 17: return x * y;
    case 18: return y * z;
    case 19: return x * z;
    default: return y * z;
  }
}

__device__ inline void dShape(int a, const double xi[3], double dN[3]) {
  double f[3], s[3];
  for (int k = 0; k < 3; ++k) {
    const bool hi = (a >> k) & 1;
    f[k]          = hi ? xi[k] : 1.0 - xi[k];
    s[k]          = hi ? 1.0 : -1.0;
  }
  dN[0] = s[0] * f[1] * f[2];
  dN[1] = f[0] * s[1] * f[2];
  dN[2] = f[0] * f[1] * s[2];
}

__device__ inline double det3(const double A[3][3]) {
  return A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
         A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
}

__device__ inline void jacobianAt(const double (*X)[3], const double xi[3], double J[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) J[i][j] = 0.0;
  for (int a = 0; a < 8; ++a) {
    double dN[3];
    dShape(a, xi, dN);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) J[i][j] += X[a][i] * dN[j];
  }
}

struct EasLds {
  double X[8][3], U[8][3];
  double alpha[NE];
  double T0[6][6];     // (T(xi0) detJ0)^-1
  double G[8][8][3];   // [gp][node][dir]
  double F[8][9];
  double S[8][6];      // S dV
  double C[8][21];     // C dV (packed)
  double f[8][NE];     // E21 mode scalings f_c(xi_gp) / detJ_gp: column c of M = T0^-1[:, kRow[c]] f_c
  double B[6][24];     // current Gauss point
  double CB[6][24];    // C dV B
  double CT[6][6];     // C dV T0^-1
  double Ch[6][6];     // T0^-T C dV T0^-1
  double Z[6][24];     // (C dV T0^-1)^T B
  double TS[6];        // T0^-T S dV
  double A[NE][NE];    // D, then its LDL^T factors
  double Lm[NE][24];   // L, then D^{-1} L
  double rt[NE];       // Rt, then D^{-1} Rt
  double K[36][9];     // K_uu upper blocks
  double Ru[24];
  int bad;
};

template <int KIND>
__device__ inline bool material(double lam, double mu, const double Ev[6], double S[6], double Cp[21]) {
  if constexpr (KIND == IK_SVK) {
    const double tr = Ev[0] + Ev[1] + Ev[2];
    for (int i = 0; i < 3; ++i) S[i] = lam * tr + 2.0 * mu * Ev[i];
    for (int i = 3; i < 6; ++i) S[i] = mu * Ev[i];
    for (int i = 0; i < 21; ++i) Cp[i] = 0.0;
    for (int p = 0; p < 3; ++p)
      for (int q = p; q < 3; ++q) Cp[sym6(p, q)] = lam + (p == q ? 2.0 * mu : 0.0);
    for (int p = 3; p < 6; ++p) Cp[sym6(p, p)] = mu;
    return true;
  } else {
    double C[3][3];
    for (int v = 0; v < 6; ++v) {
      const int i = vI(v), j = vJ(v);
      const double e = v < 3 ? Ev[v] : 0.5 * Ev[v];
      C[i][j] = C[j][i] = 2.0 * e + (i == j ? 1.0 : 0.0);
    }
    const double detC = det3(C);
    if (!(detC > 0.0)) return false;
    const double id = 1.0 / detC;
    double Ci[3][3];
    Ci[0][0] = (C[1][1] * C[2][2] - C[1][2] * C[2][1]) * id;
    Ci[0][1] = (C[0][2] * C[2][1] - C[0][1] * C[2][2]) * id;
    Ci[0][2] = (C[0][1] * C[1][2] - C[0][2] * C[1][1]) * id;
    Ci[1][0] = (C[1][2] * C[2][0] - C[1][0] * C[2][2]) * id;
    Ci[1][1] = (C[0][0] * C[2][2] - C[0][2] * C[2][0]) * id;
    Ci[1][2] = (C[0][2] * C[1][0] - C[0][0] * C[1][2]) * id;
    Ci[2][0] = (C[1][0] * C[2][1] - C[1][1] * C[2][0]) * id;
    Ci[2][1] = (C[0][1] * C[2][0] - C[0][0] * C[2][1]) * id;
    Ci[2][2] = (C[0][0] * C[1][1] - C[0][1] * C[1][0]) * id;
    const double lnJ = log(sqrt(detC));
    const double c2  = mu - lam * lnJ;
    for (int v = 0; v < 6; ++v) {
      const double ci = Ci[vI(v)][vJ(v)];
      S[v]            = mu * ((v < 3 ? 1.0 : 0.0) - ci) + lam * lnJ * ci;
    }
    for (int p = 0; p < 6; ++p)
      for (int q = p; q < 6; ++q) {
        const int I = vI(p), J = vJ(p), K = vI(q), L = vJ(q);
        Cp[sym6(p, q)] = lam * Ci[I][J] * Ci[K][L] + c2 * (Ci[I][K] * Ci[J][L] + Ci[I][L] * Ci[J][K]);
      }
    return true;
  }
}

template <int KIND, bool WANT_K>
__global__ __launch_bounds__(64) void easElement(const double* __restrict__ X, const double* __restrict__ U,
                                                 const int32_t* __restrict__ conn,
                                                 const double* __restrict__ alpha, int64_t e0, double lam,
                                                 double mu, const int32_t* __restrict__ dest,
                                                 const int32_t* __restrict__ rdest, double* __restrict__ Kebe,
                                                 double* __restrict__ Rebe, const uint8_t* __restrict__ fbits,
                                                 double3 trac, double* __restrict__ cache,
                                                 unsigned long long* __restrict__ err) {
  __shared__ EasLds s;
  const int t     = threadIdx.x;
  const int64_t e = e0 + blockIdx.x;
  if (t < 8) {
    const int64_t nd = conn[e * 8 + t];
    for (int k = 0; k < 3; ++k) {
      s.X[t][k] = X[nd * 3 + k];
      s.U[t][k] = U[nd * 3 + k];
    }
  }
  if (t < NE) s.alpha[t] = alpha[e * NE + t];
  if (t == 0) s.bad = 0;
  __syncthreads();
  // (T(xi0) detJ0)^-1 (tensorutils.hh:440-465; helperfunctions.hh:19-25). The Voigt strain transformation is
  // a representation of the Jacobian, T(J)^-1 = T(J^-1), so each of 36 lanes writes one entry directly.
  if (t < 36) {
    const double xc[3] = {0.5, 0.5, 0.5};
    double J[3][3];
    jacobianAt(s.X, xc, J);
    const double dJ = det3(J);
    const double id = 1.0 / dJ;
    double Ji[3][3];
    Ji[0][0]       = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
    Ji[0][1]       = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    Ji[0][2]       = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    Ji[1][0]       = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
    Ji[1][1]       = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[1][2]       = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[2][0]       = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
    Ji[2][1]       = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[2][2]       = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    const int r = t / 6, c = t % 6;
    const int a = vI(r), b = vJ(r), k = vI(c), l = vJ(c);
    const double v = Ji[k][a] * Ji[l][b] + Ji[l][a] * Ji[k][b];
    s.T0[r][c]     = (a == b ? 0.5 * v : v) / fabs(dJ);
  }
  __syncthreads();
  // Gauss point kinematics (lanes 0-7)
  if (t < 8) {
    const double d = 0.28867513459481288225;
    double xi[3];
    for (int k = 0; k < 3; ++k) xi[k] = ((t >> k) & 1) ? 0.5 + d : 0.5 - d;
    double dN[8][3];
    for (int a = 0; a < 8; ++a) dShape(a, xi, dN[a]);
    double J[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double acc = 0.0;
        for (int a = 0; a < 8; ++a) acc += s.X[a][i] * dN[a][j];
        J[i][j] = acc;
      }
    const double detJ = det3(J);
    const double id   = 1.0 / detJ;
    double Ji[3][3];
    Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    double F[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int a = 0; a < 8; ++a) {
      double g[3];
      for (int i = 0; i < 3; ++i) g[i] = Ji[0][i] * dN[a][0] + Ji[1][i] * dN[a][1] + Ji[2][i] * dN[a][2];
      for (int i = 0; i < 3; ++i) {
        s.G[t][a][i] = g[i];
        for (int j = 0; j < 3; ++j) F[i][j] += s.U[a][i] * g[j];
      }
    }
    double Ev[6];
    for (int v = 0; v < 6; ++v) {
      const int I = vI(v), Jj = vJ(v);
      const double c = F[0][I] * F[0][Jj] + F[1][I] * F[1][Jj] + F[2][I] * F[2][Jj];
      Ev[v]          = v < 3 ? 0.5 * (c - 1.0) : c;  // engineering shear 2E_IJ = C_IJ
    }
    // M = T0^{-1} Mtilde / detJ; E_enh = E + M alpha
    const double x = 2 * xi[0] - 1.0, y = 2 * xi[1] - 1.0, z = 2 * xi[2] - 1.0;
    const double sc = 1.0 / fabs(detJ);
    for (int c = 0; c < NE; ++c) {
      const double f = ansatz(c, x, y, z) * sc;
      const int rr   = kRow[c];
      s.f[t][c] = f;
      for (int r = 0; r < 6; ++r) Ev[r] += s.T0[r][rr] * f * s.alpha[c];
    }
    double Sv[6], Cp[21];
    if (!material<KIND>(lam, mu, Ev, Sv, Cp)) {
      atomicMin(err, (unsigned long long)e);
      s.bad = 1;
    }
    const double dV = fabs(detJ) * 0.125;
    for (int i = 0; i < 9; ++i) s.F[t][i] = F[i / 3][i % 3];
    for (int v = 0; v < 6; ++v) s.S[t][v] = Sv[v] * dV;
    for (int i = 0; i < 21; ++i) s.C[t][i] = Cp[i] * dV;
  }
  __syncthreads();
  if (s.bad) return;  // uniform: the error flag is already set

  // accumulators: lane t < 36 owns K_uu block t; lanes own D / L / Rt / R_u entries (strided). The entry
  // each lane owns does not change over the Gauss points, so its indices are decoded once here.
  double kb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double dacc[4] = {0, 0, 0, 0};   // D upper entries t, t+64, t+128, t+192 (< 231)
  double lacc[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // L entries t + 64 j (< 504)
  double racc = 0.0;               // lanes 0-20: Rt ; lanes 32-55: R_u
  int blkA = 0, blkB = 0;
  if (t < 36) {
    int r = t, a = 0;
    while (r >= 8 - a) {
      r -= 8 - a;
      ++a;
    }
    blkA = a;
    blkB = a + r;
  }
  int dI[4], dC[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int r = t + 64 * j, i = 0;
    if (r < 231)
      while (r >= NE - i) {
        r -= NE - i;
        ++i;
      }
    dI[j] = i;
    dC[j] = i + r;
  }
  int dKi[4], dKc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    dKi[j] = kRow[dI[j] < NE ? dI[j] : 0];
    dKc[j] = kRow[dC[j] < NE ? dC[j] : 0];
  }
  int lI[8], lCol[8], lK[8];  // L entry idx = t + 64 j: row (mode) i, column, kRow[i]
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int idx = t + 64 * j;
    lI[j]   = idx / 24;
    lCol[j] = idx % 24;
    lK[j]   = kRow[lI[j] < NE ? lI[j] : 0];
  }
  const int rK = kRow[t < NE ? t : 0];
  int bV[3], bCol[3], bG[3], bK[3];  // B build / C B / Z: entry idx = t + 64 j of the 6 x 24 matrices
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int idx = t + 64 * j;
    bV[j]   = idx / 24;
    bCol[j] = idx % 24;
    bG[j]   = bCol[j] / 3;
    bK[j]   = bCol[j] % 3;
  }
  for (int g = 0; g < 8; ++g) {
    // B (6x24) and CT = C dV T0^-1 for this Gauss point
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if (t + 64 * j < 144) {
        const double* F  = s.F[g];
        const double* ga = s.G[g][bG[j]];
        const int I = vI(bV[j]), J = vJ(bV[j]), k = bK[j];
        double bv = F[k * 3 + I] * ga[J];
        if (I != J) bv += F[k * 3 + J] * ga[I];
        s.B[bV[j]][bCol[j]] = bv;
      }
    }
    if (t < 36) {
      const int v = t / 6, w = t % 6;
      double acc = 0.0;
#pragma unroll
      for (int u = 0; u < 6; ++u) acc += s.C[g][sym6(v, u)] * s.T0[u][w];
      s.CT[v][w] = acc;
    }
    __syncthreads();
    // C B (for K_uu), Z = CT^T B (for L), Ch = T0^-T CT (for D), TS = T0^-T S (for Rt)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if (t + 64 * j < 144) {
        double acc = 0.0, accz = 0.0;
#pragma unroll
        for (int w = 0; w < 6; ++w) {
          if constexpr (WANT_K) acc += s.C[g][sym6(bV[j], w)] * s.B[w][bCol[j]];
          accz += s.CT[w][bV[j]] * s.B[w][bCol[j]];
        }
        if constexpr (WANT_K) s.CB[bV[j]][bCol[j]] = acc;
        s.Z[bV[j]][bCol[j]] = accz;
      }
    }
    if (t < 36) {
      const int p = t / 6, q = t % 6;
      double acc = 0.0;
#pragma unroll
      for (int v = 0; v < 6; ++v) acc += s.T0[v][p] * s.CT[v][q];
      s.Ch[p][q] = acc;
    } else if (t >= 40 && t < 46) {
      const int w = t - 40;
      double acc = 0.0;
#pragma unroll
      for (int v = 0; v < 6; ++v) acc += s.T0[v][w] * s.S[g][v];
      s.TS[w] = acc;
    }
    __syncthreads();
    if constexpr (WANT_K) {
      if (t < 36) {
        // K_uu block (a, b): sum_v B[v][3a+k] (C B)[v][3b+l] + g_a^T S g_b delta_kl
        const double* ga = s.G[g][blkA];
        const double* gb = s.G[g][blkB];
        const double* Sv = s.S[g];
        const double kg  = ga[0] * (Sv[0] * gb[0] + Sv[5] * gb[1] + Sv[4] * gb[2]) +
                          ga[1] * (Sv[5] * gb[0] + Sv[1] * gb[1] + Sv[3] * gb[2]) +
                          ga[2] * (Sv[4] * gb[0] + Sv[3] * gb[1] + Sv[2] * gb[2]);
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int l = 0; l < 3; ++l) {
            double acc = 0.0;
#pragma unroll
            for (int v = 0; v < 6; ++v) acc += s.B[v][3 * blkA + k] * s.CB[v][3 * blkB + l];
            kb[k * 3 + l] += acc + (k == l ? kg : 0.0);
          }
      }
    }
    // D[i][c] += f_i f_c Ch[kRow_i][kRow_c]; L[i][col] += f_i Z[kRow_i][col]; Rt[i] += f_i TS[kRow_i]
    const double* fg = s.f[g];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (t + 64 * j < 231) dacc[j] += fg[dI[j]] * fg[dC[j]] * s.Ch[dKi[j]][dKc[j]];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (t + 64 * j < NE * 24) lacc[j] += fg[lI[j]] * s.Z[lK[j]][lCol[j]];
    if (t < NE) {
      racc += fg[t] * s.TS[rK];
    } else if (t >= 32 && t < 56) {
      const int col = t - 32;
      double acc    = 0.0;
#pragma unroll
      for (int v = 0; v < 6; ++v) acc += s.B[v][col] * s.S[g][v];
      racc += acc;
    }
    __syncthreads();
  }
  // spill accumulators to LDS
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (t + 64 * j < 231) {
      s.A[dI[j]][dC[j]] = dacc[j];
      s.A[dC[j]][dI[j]] = dacc[j];
    }
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int idx = t + 64 * j;
    if (idx < NE * 24) s.Lm[idx / 24][idx % 24] = lacc[j];
  }
  if (t < NE) s.rt[t] = racc;
  if (t >= 32 && t < 56) s.Ru[t - 32] = racc;
  if constexpr (WANT_K) {
    if (t < 36)
      for (int i = 0; i < 9; ++i) s.K[t][i] = kb[i];
  }
  __syncthreads();

  // Y = D^-1 [L | Rt] by Gauss-Jordan elimination on the augmented 21 x 46 matrix (D is SPD: no pivoting,
  // like the LDL^T it replaces; enhancedassumedstrains.hh inverts D, :243,294,344). Lane c < 46 keeps column
  // c in registers; each step the pivot column is broadcast through LDS.
  double col[NE];
#pragma unroll
  for (int i = 0; i < NE; ++i) col[i] = t < NE ? s.A[i][t] : (t < NE + 24 ? s.Lm[i][t - NE] : (t == NE + 24 ? s.rt[i] : 0.0));
  __syncthreads();
  double* pivbuf = &s.A[0][0];  // D is in the lanes' registers from here on
#pragma unroll
  for (int j = 0; j < NE; ++j) {
    if (t == j)
#pragma unroll
      for (int i = 0; i < NE; ++i) pivbuf[i] = col[i];
    __syncthreads();
    const double f = col[j] / pivbuf[j];
#pragma unroll
    for (int i = 0; i < NE; ++i)
      if (i != j) col[i] -= pivbuf[i] * f;
    col[j] = f;
    __syncthreads();
  }
  // lanes NE..NE+23 hold Y = D^-1 L (column c = t - NE), lane NE+24 holds y = D^-1 Rt; cache them for the
  // alpha update (layout per element: Y 21 x 24 row-major, then y)
  double* cy = cache + e * (NE * 24 + NE);
  if (t >= NE && t < NE + 24) {
#pragma unroll
    for (int i = 0; i < NE; ++i) cy[i * 24 + (t - NE)] = col[i];
  } else if (t == NE + 24) {
#pragma unroll
    for (int i = 0; i < NE; ++i) {
      cy[NE * 24 + i] = col[i];
      s.rt[i]         = col[i];  // y for the residual lanes
    }
  }
  // K = K_uu - L^T Y on the upper blocks: lane of column c = 3b+l computes rows 3a+k of blocks a < b and,
  // for the diagonal block, k <= l (mirrored afterwards: enhancedassumedstrains.hh:291-296)
  if constexpr (WANT_K) {
    if (t >= NE && t < NE + 24) {
      const int c = t - NE, b = c / 3, l = c % 3;
      for (int r = 0; r <= c; ++r) {
        const int a = r / 3, k = r % 3;
        if (a == b && k > l) continue;
        double acc = s.K[upperBlock(a, b)][3 * k + l];
#pragma unroll
        for (int i = 0; i < NE; ++i) acc -= s.Lm[i][r] * col[i];
        s.K[upperBlock(a, b)][3 * k + l] = acc;
        if (a == b && k != l) s.K[upperBlock(a, b)][3 * l + k] = acc;
      }
    }
  }
  __syncthreads();
  if constexpr (WANT_K) {
    if (t < 36) {
      const double* out = s.K[t];
      if (dest) {
        const int32_t code = dest[e * 36 + t];
        double* dst        = Kebe + (int64_t)(code >> 1) * 9;
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) dst[3 * k + l] = (code & 1) ? out[3 * l + k] : out[3 * k + l];
      } else {
        double* dst = Kebe + (e * 36 + t) * 9;
        for (int i = 0; i < 9; ++i) dst[i] = out[i];
      }
    }
  }
  if (t >= 40 && t < 64) {
    const int r = t - 40;
    double acc  = s.Ru[r];
    for (int i = 0; i < NE; ++i) acc -= s.Lm[i][r] * s.rt[i];
    const unsigned bits = fbits ? fbits[e] : 0u;
    if (bits) {  // Traction skill, after the EAS element's condensed internal forces
      const int k = r % 3;
      acc += faceLoad(X, conn, e, bits, r / 3, k == 0 ? trac.x : (k == 1 ? trac.y : trac.z));
    }
    Rebe[rdest ? (int64_t)rdest[e * 8 + r / 3] * 3 + r % 3 : e * 24 + r] = acc;
  }
}

// alpha_e -= y_e + Y_e du_e (thread per (element, mode))
__global__ __launch_bounds__(256) void easUpdateKernel(const double* __restrict__ cache,
                                                       const int32_t* __restrict__ conn,
                                                       const double* __restrict__ du, int64_t ne,
                                                       double* __restrict__ alpha) {
  const int64_t tid = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (tid >= ne * NE) return;
  const int64_t e = tid / NE;
  const int i     = (int)(tid - e * NE);
  const double* cy = cache + e * (NE * 24 + NE);
  double acc       = cy[NE * 24 + i];
  for (int a = 0; a < 8; ++a) {
    const int64_t nd = conn[e * 8 + a];
    for (int k = 0; k < 3; ++k) acc += cy[i * 24 + a * 3 + k] * du[nd * 3 + k];
  }
  alpha[tid] -= acc;
}

template <bool WK>
void launchEas(ik_handle* h, int64_t e0, int64_t count, const int32_t* dest, const int32_t* rdest, double* Kebe) {
  if (h->material == IK_SVK)
    hipLaunchKernelGGL((easElement<IK_SVK, WK>), dim3((unsigned)count), dim3(64), 0, h->stream, h->X.p, h->u.p,
                       h->conn.p, h->alpha.p, e0, h->lam, h->mu, dest, rdest, Kebe, h->ebeR.p, h->faceBits.p, loadedTraction(h), h->easCache.p,
                       h->errFlag.p);
  else
    hipLaunchKernelGGL((easElement<IK_NEOHOOKE, WK>), dim3((unsigned)count), dim3(64), 0, h->stream, h->X.p,
                       h->u.p, h->conn.p, h->alpha.p, e0, h->lam, h->mu, dest, rdest, Kebe, h->ebeR.p, h->faceBits.p, loadedTraction(h), h->easCache.p,
                       h->errFlag.p);
}

}  // namespace

void easStage(ik_handle* h, int64_t e0, int64_t count, bool wantK, const int32_t* dest, const int32_t* rdest) {
  h->easCache.ensure((size_t)h->ne * (NE * 24 + NE));
  if (wantK)
    launchEas<true>(h, e0, count, dest, rdest, h->ebeK64.p);
  else
    launchEas<false>(h, e0, count, nullptr, rdest, nullptr);
}

void easUpdate(ik_handle* h, const double* duDev) {
  require(h->easCacheValid, IK_EINVAL, "EAS update needs D, L, Rt of the current state (assemble first)");
  const int64_t n = h->ne * NE;
  hipLaunchKernelGGL(easUpdateKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->easCache.p,
                     h->conn.p, duDev, h->ne, h->alpha.p);
  h->easCacheValid = false;
}

}  // namespace ik
