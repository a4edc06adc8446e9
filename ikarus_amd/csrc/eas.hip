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
  double T0[6][6];     // T(xi0) detJ0, inverted in place
  double G[8][8][3];   // [gp][node][dir]
  double F[8][9];
  double S[8][6];      // S dV
  double C[8][21];     // C dV (packed)
  double M[8][6][NE];  // M(xi_gp)
  double B[6][24];     // current Gauss point
  double CB[6][24];    // C dV B
  double CM[6][NE];    // C dV M
  double A[NE][NE];    // D, then its LDL^T factors
  double Lm[NE][24];   // L, then D^{-1} L
  double rt[NE];       // Rt, then D^{-1} Rt
  double K[36][9];     // K_uu upper blocks
  double Ru[24];
  double Y[NE][25];    // D^{-1} [L | Rt]
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
  // T(xi0) * detJ0 (tensorutils.hh:440-465; helperfunctions.hh:19-25), then inverted by lane 0
  if (t < 36) {
    const double xc[3] = {0.5, 0.5, 0.5};
    double J[3][3];
    jacobianAt(s.X, xc, J);
    const double dJ = fabs(det3(J));
    const int r = t / 6, c = t % 6;
    const int a = vI(r), b = vJ(r), k = vI(c), l = vJ(c);
    // Jt[i][j] = J[j][i]
    const double v = J[k][a] * J[l][b] + J[l][a] * J[k][b];
    s.T0[r][c]     = (a == b ? 0.5 * v : v) * dJ;
  }
  __syncthreads();
  if (t == 0) {
    // Gauss-Jordan with partial pivoting on the 6x6 (in LDS, into A's first 36 slots as scratch)
    double Mx[6][12];
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 12; ++c) Mx[r][c] = c < 6 ? s.T0[r][c] : (c - 6 == r ? 1.0 : 0.0);
    for (int c = 0; c < 6; ++c) {
      int p = c;
      for (int r = c + 1; r < 6; ++r)
        if (fabs(Mx[r][c]) > fabs(Mx[p][c])) p = r;
      if (p != c)
        for (int k = 0; k < 12; ++k) {
          const double tmp = Mx[p][k];
          Mx[p][k]         = Mx[c][k];
          Mx[c][k]         = tmp;
        }
      const double piv = 1.0 / Mx[c][c];
      for (int k = 0; k < 12; ++k) Mx[c][k] *= piv;
      for (int r = 0; r < 6; ++r) {
        if (r == c) continue;
        const double f = Mx[r][c];
        for (int k = 0; k < 12; ++k) Mx[r][k] -= f * Mx[c][k];
      }
    }
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) s.T0[r][c] = Mx[r][c + 6];
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
      for (int r = 0; r < 6; ++r) {
        const double m = s.T0[r][rr] * f;
        s.M[t][r][c]   = m;
        Ev[r] += m * s.alpha[c];
      }
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

  // accumulators: lane t < 36 owns K_uu block t; lanes own D / L / Rt / R_u entries (strided)
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
  for (int g = 0; g < 8; ++g) {
    // B (6x24), C B and C M for this Gauss point
    for (int idx = t; idx < 144; idx += 64) {
      const int v = idx / 24, col = idx % 24, a = col / 3, k = col % 3;
      const double* F  = s.F[g];
      const double* ga = s.G[g][a];
      const int I = vI(v), J = vJ(v);
      double bv = F[k * 3 + I] * ga[J];
      if (I != J) bv += F[k * 3 + J] * ga[I];
      s.B[v][col] = bv;
    }
    for (int idx = t; idx < 6 * NE; idx += 64) {
      const int v = idx / NE, c = idx % NE;
      double acc = 0.0;
      for (int w = 0; w < 6; ++w) acc += s.C[g][sym6(v, w)] * s.M[g][w][c];
      s.CM[v][c] = acc;
    }
    __syncthreads();
    if constexpr (WANT_K) {
      for (int idx = t; idx < 144; idx += 64) {
        const int v = idx / 24, col = idx % 24;
        double acc = 0.0;
        for (int w = 0; w < 6; ++w) acc += s.C[g][sym6(v, w)] * s.B[w][col];
        s.CB[v][col] = acc;
      }
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
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) {
            double acc = 0.0;
            for (int v = 0; v < 6; ++v) acc += s.B[v][3 * blkA + k] * s.CB[v][3 * blkB + l];
            kb[k * 3 + l] += acc + (k == l ? kg : 0.0);
          }
      }
    }
    // D upper (231 entries) and L (504 entries)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int idx = t + 64 * j;
      if (idx < 231) {
        int r = idx, i = 0;
        while (r >= NE - i) {
          r -= NE - i;
          ++i;
        }
        const int c = i + r;
        double acc  = 0.0;
        for (int v = 0; v < 6; ++v) acc += s.M[g][v][i] * s.CM[v][c];
        dacc[j] += acc;
      }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int idx = t + 64 * j;
      if (idx < NE * 24) {
        const int i = idx / 24, col = idx % 24;
        double acc  = 0.0;
        for (int v = 0; v < 6; ++v) acc += s.CM[v][i] * s.B[v][col];
        lacc[j] += acc;
      }
    }
    if (t < NE) {
      double acc = 0.0;
      for (int v = 0; v < 6; ++v) acc += s.M[g][v][t] * s.S[g][v];
      racc += acc;
    } else if (t >= 32 && t < 56) {
      const int col = t - 32;
      double acc    = 0.0;
      for (int v = 0; v < 6; ++v) acc += s.B[v][col] * s.S[g][v];
      racc += acc;
    }
    __syncthreads();
  }
  // spill accumulators to LDS
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int idx = t + 64 * j;
    if (idx < 231) {
      int r = idx, i = 0;
      while (r >= NE - i) {
        r -= NE - i;
        ++i;
      }
      const int c = i + r;
      s.A[i][c]   = dacc[j];
      s.A[c][i]   = dacc[j];
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
  // LDL^T of D, right-looking, lower triangle in place: A[i][j] (i > j) = l_ij, A[j][j] = d_j
  for (int j = 0; j < NE; ++j) {
    const double dj = s.A[j][j];
    const int m     = NE - 1 - j;  // trailing size
    for (int idx = t; idx < m * (m + 1) / 2; idx += 64) {
      int r = idx, ii = 0;
      while (r > ii) {
        r -= ii + 1;
        ++ii;
      }
      const int i = j + 1 + ii, k = j + 1 + r;  // k <= i
      s.A[i][k] -= s.A[i][j] * s.A[k][j] / dj;
    }
    __syncthreads();
    if (t > j && t < NE) s.A[t][j] /= dj;
    __syncthreads();
  }
  // Y = D^{-1} [L | Rt]: lane c < 25 solves column c
  if (t < 25) {
    double y[NE];
    for (int i = 0; i < NE; ++i) y[i] = t < 24 ? s.Lm[i][t] : s.rt[i];
    for (int i = 0; i < NE; ++i)
      for (int k = 0; k < i; ++k) y[i] -= s.A[i][k] * y[k];
    for (int i = 0; i < NE; ++i) y[i] /= s.A[i][i];
    for (int i = NE - 1; i >= 0; --i)
      for (int k = i + 1; k < NE; ++k) y[i] -= s.A[k][i] * y[k];
    // cache layout per element: Y (21 x 24) row-major, then y (21)
    double* cy = cache + e * (NE * 24 + NE);
    for (int i = 0; i < NE; ++i) {
      s.Y[i][t] = y[i];
      if (t < 24)
        cy[i * 24 + t] = y[i];
      else
        cy[NE * 24 + i] = y[i];
    }
  }
  __syncthreads();
  // K = K_uu - L^T Y (upper blocks), R = R_u - L^T y
  if constexpr (WANT_K) {
    if (t < 36) {
      double* dst = Kebe + (e * 36 + t) * 9;
      double out[9];
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          double acc = s.K[t][k * 3 + l];
          const int r = 3 * blkA + k, c = 3 * blkB + l;
          for (int i = 0; i < NE; ++i) acc -= s.Lm[i][r] * s.Y[i][c];
          out[k * 3 + l] = acc;
        }
      if (blkA == blkB) {  // mirror the diagonal block's upper triangle (enhancedassumedstrains.hh:291-296)
        out[3] = out[1];
        out[6] = out[2];
        out[7] = out[5];
      }
      if (dest) {
        const int32_t code = dest[e * 36 + t];
        dst                = Kebe + (int64_t)(code >> 1) * 9;
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) dst[3 * k + l] = (code & 1) ? out[3 * l + k] : out[3 * k + l];
      } else {
        for (int i = 0; i < 9; ++i) dst[i] = out[i];
      }
    }
  }
  if (t >= 40 && t < 64) {
    const int r = t - 40;
    double acc  = s.Ru[r];
    for (int i = 0; i < NE; ++i) acc -= s.Lm[i][r] * s.Y[i][24];
    const unsigned bits = fbits ? fbits[e] : 0u;
    if (bits) {  // Traction skill, after the EAS element's condensed internal forces
      const double tk[3] = {trac.x, trac.y, trac.z};
      acc += faceLoad(X, conn, e, bits, r / 3, tk[r % 3]);
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
