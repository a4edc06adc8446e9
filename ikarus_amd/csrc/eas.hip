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
// One 64-lane workgroup (one wave) per element, all lanes busy in every phase: Gauss-point kinematics and
// material per (point, node) lane pair, K_uu in the Ci form per (node, node) lane pair, D / L / Rt / R_u by
// owned entries, forward elimination of [D | L | Rt] one column per lane (D = L_D diag(d) L_D^T), and the
// rank-21 update of K and R. Outputs go to the same element-by-element staging as the plain element (36 upper
// blocks + 24 residual entries), plus the cache of the alpha update: diag(d)^{-1} L_D^{-1} L (21x24),
// diag(d)^{-1} L_D^{-1} Rt (21) and L_D (210), so that D^{-1} v = L_D^{-T} diag(d)^{-1} L_D^{-1} v.
#include "ik_internal.hh"

namespace ik {
namespace {

constexpr int NE = 21;  // enhanced modes
// doubles per element of the alpha-update cache: What (24 columns of 21), what (21), the strictly lower L_D (210)
constexpr int kEasCache = NE * 24 + NE + NE * (NE - 1) / 2;

__host__ __device__ constexpr int vI(int v) { return v < 3 ? v : (v == 3 ? 1 : 0); }
__host__ __device__ constexpr int vJ(int v) { return v < 3 ? v : (v == 5 ? 1 : 2); }
__host__ __device__ constexpr int sym6(int p, int q) {
  return p <= q ? p * 6 - (p * (p - 1)) / 2 + (q - p) : q * 6 - (q * (q - 1)) / 2 + (p - q);
}
__host__ __device__ constexpr int upperBlock(int a, int b) { return 8 * a - (a * (a - 1)) / 2 + (b - a); }

// Row-major upper triangle of an n x n matrix: entry idx in [0, n(n+1)/2) -> (row i, column c >= i). Rows
// before i hold S(i) = i n - i(i-1)/2 entries, so i is the root of that quadratic, rounded down (the float
// root is off by at most one, which the integer checks correct).
__device__ inline void triDecode(int idx, int n, int& i, int& c) {
  const float b = 2.0f * n + 1.0f;
  int r         = (int)((b - sqrtf(fmaxf(b * b - 8.0f * idx, 0.0f))) * 0.5f);
  r             = r < 0 ? 0 : (r > n - 1 ? n - 1 : r);
  auto S        = [n](int k) { return k * n - (k * (k - 1)) / 2; };
  if (r < n - 1 && S(r + 1) <= idx) ++r;
  if (r > 0 && S(r) > idx) --r;
  i = r;
  c = r + (idx - S(r));
}

// E21 ansatz in reference coordinates: column c has one nonzero, in Voigt row kRow[c], of value
// f_c(x, y, z) with x = 2 xi - 1 etc. (linearandglstrains.hh:300-322)
__constant__ int kRow[NE] = {0, 1, 2, 3, 3, 4, 4, 5, 5, 3, 3, 4, 4, 5, 5, 0, 0, 1, 1, 2, 2};
__host__ __device__ constexpr int kRowOf(int c) {
  return c < 3 ? c : (c < 9 ? 3 + (c - 3) / 2 : (c < 15 ? 3 + (c - 9) / 2 : (c - 15) / 2));
}
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

// value of lane `lane` (uniform) of a 64-bit register, as two 32-bit lane reads into scalar registers
__device__ inline double readLane(double x, int lane) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo          = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
  const unsigned hi          = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// maximum over the 64 lanes of the wavefront (every lane active), as a uniform value
__device__ inline double waveMax(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
  return readLane(x, 0);
}

// Voigt index of the symmetric pair (i, j): 00 11 22 12 02 01
__host__ __device__ constexpr int v6(int i, int j) {
  return i == j ? i : (i + j == 3 ? 3 : (i + j == 2 ? 4 : 5));
}

// One 64-lane workgroup (one wave) per element. The LDS (13.6 KB; 11.2 KB for the residual alone) holds two
// unions: the K_uu inputs (later the parked K_uu blocks) and the point data of
// the D / L / Rt loop (later the condensation operands).
template <bool WANT_K>
struct EasLds {
  static constexpr int NK = WANT_K ? 8 : 1;  // the residual-only kernel keeps no K_uu inputs
  double X[8][3], U[8][3];
  double alpha[NE];
  double T0[6][6];  // (T(xi0) detJ0)^-1
  union {
    struct {
      double G[8][8][3];  // [gp][node][dir] shape-function gradients g_a
      double S[8][6];     // S dV
      double P[NK][8][3]; // F Ci g_a
      double Ci[NK][6];   // Ci (Voigt order), the identity for SVK
      double Phi[NK][6];  // F Ci F^T (Voigt order)
      double w[NK][2];    // lambda dV, c2 dV: material tangent weights of the Ci form
    } k;
    struct {
      double Kuu[WANT_K ? 36 : 1][9];  // K_uu upper blocks while the elimination runs
    } q;
  } uk;
  union {
    struct {
      double F[8][9];
      double C[8][21];    // C dV (packed)
      double f[8][NE];    // E21 mode scalings f_c(xi_gp) / detJ_gp: column c of M = T0^-1[:, kRow[c]] f_c
      double CT[6][6];    // current point: C dV T0^-1
      double Ch[6][6];    // T0^-T C dV T0^-1
      double TS[6];       // T0^-T S dV
      double Z[6][24];    // (C dV T0^-1)^T B
      double ru[8][8][3]; // R_u contribution F S dV g_a of point gp, node a
    } g;
    struct {
      double A[NE][24];   // D (columns 0-20), then What = diag(d)^-1 L_D^-1 L
      double Lm[NE][24];  // L, then W = L_D^-1 L (row j written at elimination step j)
      double rt[NE];      // Rt, then what = diag(d)^-1 L_D^-1 Rt
      double Ru[24];
    } c;
  } u;
  int bad;
};

// S and C (Voigt) at the enhanced strain, plus the factors of the tangent's Ci form
// C = lambda Ci(x)Ci + c2 (Ci box Ci + transposed): Ci (Voigt) and c2 (Neo-Hooke: mu - lambda ln J; SVK: Ci = I,
// c2 = mu)
template <int KIND>
__device__ inline bool material(double lam, double mu, const double Ev[6], double S[6], double Cp[21], double Ci6[6],
                                double& c2) {
  if constexpr (KIND == IK_SVK) {
    const double tr = Ev[0] + Ev[1] + Ev[2];
    for (int i = 0; i < 3; ++i) S[i] = lam * tr + 2.0 * mu * Ev[i];
    for (int i = 3; i < 6; ++i) S[i] = mu * Ev[i];
    for (int i = 0; i < 21; ++i) Cp[i] = 0.0;
    for (int p = 0; p < 3; ++p)
      for (int q = p; q < 3; ++q) Cp[sym6(p, q)] = lam + (p == q ? 2.0 * mu : 0.0);
    for (int p = 3; p < 6; ++p) Cp[sym6(p, p)] = mu;
    for (int v = 0; v < 6; ++v) Ci6[v] = v < 3 ? 1.0 : 0.0;
    c2 = mu;
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
    c2               = mu - lam * lnJ;
    for (int v = 0; v < 6; ++v) {
      const double ci = Ci[vI(v)][vJ(v)];
      Ci6[v]          = ci;
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

// y = A x for a symmetric 3x3 A in Voigt order
__device__ inline void symMul(const double* A6, const double x[3], double y[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) y[i] = A6[v6(i, 0)] * x[0] + A6[v6(i, 1)] * x[1] + A6[v6(i, 2)] * x[2];
}

// Phases (all 64 lanes): (1) Gauss-point kinematics, enhanced strain and material, lane (gp, node): each lane
// keeps its node's g_a and F Ci g_a, node 0 the point's data; (2) the point loop: K_uu in the Ci form,
//   K_ab[k][l] += lambda dV p_ak p_bl + c2 dV (p_al p_bk + (g_a.Ci g_b) Phi_kl) + (g_a.S dV g_b) delta_kl,
//   p_a = F Ci g_a, Phi = F Ci F^T (the material part of B^T C B written with C = lambda Ci(x)Ci + 2 c2 sym(Ci box
//   Ci); SVK: Ci = I, c2 = mu), lane (a, b) per 3x3 block; then D, L = sum f_i Z[kRow_i], Rt and R_u
//   = sum (F S dV g_b)_l with Z[w][3b+l] = (F Q^w g_b)_l, Q^w the symmetric tensor of column w of C dV T0^-1;
// (3) W = L_D^-1 [L | Rt] and What = diag(d)^-1 W by forward elimination, one column per lane; (4) K = K_uu -
// W^T What, lane (a, b) per block (the diagonal blocks mirrored from their upper triangle), R = R_u - W^T what.
template <int KIND, bool WANT_K>
__global__ __launch_bounds__(64) void easElement(const double* __restrict__ X, const double* __restrict__ U,
                                                 const int32_t* __restrict__ conn,
                                                 const double* __restrict__ alpha, int64_t e0, double lam,
                                                 double mu, const int32_t* __restrict__ dest,
                                                 const int32_t* __restrict__ rdest, double* __restrict__ Kebe,
                                                 double* __restrict__ Rebe, const uint8_t* __restrict__ fbits,
                                                 double3 trac, double* __restrict__ cache,
                                                 unsigned long long* __restrict__ err) {
  __shared__ EasLds<WANT_K> s;
  auto& sk        = s.uk.k;
  auto& sq        = s.uk.q;
  auto& sg        = s.u.g;
  auto& sc        = s.u.c;
  const int t     = threadIdx.x;
  const int64_t e = e0 + blockIdx.x;
  if (t < 24) {
    const int64_t nd = conn[e * 8 + t / 3];
    s.X[t / 3][t % 3] = X[nd * 3 + t % 3];
    s.U[t / 3][t % 3] = U[nd * 3 + t % 3];
  } else if (t >= 32 && t < 32 + NE) {
    s.alpha[t - 32] = alpha[e * NE + (t - 32)];
  } else if (t == 63) {
    s.bad = 0;
  }
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
  // (1) Gauss point gp = t / 8 (kinematics repeated by its 8 lanes), node na = t % 8
  {
    const int gp = t >> 3, na = t & 7;
    const double d = 0.28867513459481288225;
    double xi[3];
    for (int k = 0; k < 3; ++k) xi[k] = ((gp >> k) & 1) ? 0.5 + d : 0.5 - d;
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      double dN[3];
      dShape(a, xi, dN);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) J[i][j] += s.X[a][i] * dN[j];
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
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      double dN[3], g[3];
      dShape(a, xi, dN);
      for (int i = 0; i < 3; ++i) g[i] = Ji[0][i] * dN[0] + Ji[1][i] * dN[1] + Ji[2][i] * dN[2];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) F[i][j] += s.U[a][i] * g[j];
    }
    double go[3];
    {
      double dN[3];
      dShape(na, xi, dN);
      for (int i = 0; i < 3; ++i) go[i] = Ji[0][i] * dN[0] + Ji[1][i] * dN[1] + Ji[2][i] * dN[2];
    }
    double Ev[6];
    for (int v = 0; v < 6; ++v) {
      const int I = vI(v), Jj = vJ(v);
      const double c = F[0][I] * F[0][Jj] + F[1][I] * F[1][Jj] + F[2][I] * F[2][Jj];
      Ev[v]          = v < 3 ? 0.5 * (c - 1.0) : c;  // engineering shear 2E_IJ = C_IJ
    }
    // E_enh = E + M alpha, M = T0^{-1} Mtilde / detJ: Mtilde alpha has one term per mode in row kRow[c]
    const double x = 2 * xi[0] - 1.0, y = 2 * xi[1] - 1.0, z = 2 * xi[2] - 1.0;
    const double sc1 = 1.0 / fabs(detJ);
    double wv[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < NE; ++c) wv[kRowOf(c)] += ansatz(c, x, y, z) * s.alpha[c];
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < 6; ++q) acc += s.T0[r][q] * wv[q];
      Ev[r] += acc * sc1;
    }
    double Sv[6], Cp[21], Ci6[6], c2;
    if (!material<KIND>(lam, mu, Ev, Sv, Cp, Ci6, c2)) {
      atomicMin(err, (unsigned long long)e);
      s.bad = 1;
    }
    const double dV = fabs(detJ) * 0.125;
    double sgo[3];
    symMul(Sv, go, sgo);
    for (int i = 0; i < 3; ++i) {
      sk.G[gp][na][i]  = go[i];
      sg.ru[gp][na][i] = (F[i][0] * sgo[0] + F[i][1] * sgo[1] + F[i][2] * sgo[2]) * dV;
    }
    if constexpr (WANT_K) {
      double q[3];
      symMul(Ci6, go, q);
      for (int i = 0; i < 3; ++i) sk.P[gp][na][i] = F[i][0] * q[0] + F[i][1] * q[1] + F[i][2] * q[2];
    }
    if (na == 0) {
      for (int i = 0; i < 9; ++i) sg.F[gp][i] = F[i / 3][i % 3];
      for (int v = 0; v < 6; ++v) sk.S[gp][v] = Sv[v] * dV;
      for (int i = 0; i < 21; ++i) sg.C[gp][i] = Cp[i] * dV;
      if constexpr (WANT_K) {
        for (int v = 0; v < 6; ++v) sk.Ci[gp][v] = Ci6[v];
        // Phi = F Ci F^T
        double FC[3][3];
        for (int i = 0; i < 3; ++i) symMul(Ci6, F[i], FC[i]);  // row i of F Ci (Ci symmetric)
        for (int v = 0; v < 6; ++v) {
          const int k = vI(v), l = vJ(v);
          sk.Phi[gp][v] = FC[k][0] * F[l][0] + FC[k][1] * F[l][1] + FC[k][2] * F[l][2];
        }
        sk.w[gp][0] = lam * dV;
        sk.w[gp][1] = c2 * dV;
      }
#pragma unroll
      for (int c = 0; c < NE; ++c) sg.f[gp][c] = ansatz(c, x, y, z) * sc1;
    }
  }
  __syncthreads();
  if (s.bad) return;  // uniform: the error flag is already set

  // (2a) point loop for D, L and Rt: D upper entries t + 64 j (< 231), L columns and Rt (lanes 0-20) are owned by
  // fixed lanes; R_u (lanes 40-63) is summed from the per-point contributions afterwards
  const int ka = t >> 3, kbn = t & 7;
  double dacc[4] = {0, 0, 0, 0};
  double racc    = 0.0;
  int dI[4], dC[4], dKi[4], dKc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    triDecode(t + 64 * j < 231 ? t + 64 * j : 230, NE, dI[j], dC[j]);
    dKi[j] = kRow[dI[j]];
    dKc[j] = kRow[dC[j]];
  }
  // L: lane t < 48 owns column lc = t % 24 of rows lh + 2m (lh = t / 24)
  const int lc = t % 24, lh = t / 24;
  double lacc[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int rK = kRow[t < NE ? t : 0];
  // the one column of T0 each lane's role reads at every point: w (C T0^-1), p (T0^-T C T0^-1), w (T0^-T S)
  int chP = 0, chQ = 0;
  if (t >= 43) triDecode(t >= 48 ? t - 48 : t - 27, 6, chP, chQ);
  double t0c[6];
  {
    const int tc = t < 36 ? t % 6 : (t >= 43 ? chP : (t >= 37 ? t - 37 : 0));
#pragma unroll
    for (int u = 0; u < 6; ++u) t0c[u] = s.T0[u][tc];
  }
  for (int g = 0; g < 8; ++g) {
    if (t < 36) {
      const int v = t / 6, w = t % 6;
      double acc = 0.0;
#pragma unroll
      for (int u = 0; u < 6; ++u) acc += sg.C[g][sym6(v, u)] * t0c[u];
      sg.CT[v][w] = acc;
    }
    __syncthreads();
    // Z[w][3b+l] = (F Q^w g_b)_l (lane (w, b) < 48), Ch = T0^-T CT (lanes 48-63, then 43-47), TS (lanes 37-42)
    if (t < 48) {
      const int w = t >> 3, b = t & 7;
      double Q[6], gbv[3], h[3];
      for (int v = 0; v < 6; ++v) Q[v] = sg.CT[v][w];
      for (int i = 0; i < 3; ++i) gbv[i] = sk.G[g][b][i];
      symMul(Q, gbv, h);
#pragma unroll
      for (int l = 0; l < 3; ++l)
        sg.Z[w][3 * b + l] = sg.F[g][3 * l] * h[0] + sg.F[g][3 * l + 1] * h[1] + sg.F[g][3 * l + 2] * h[2];
    }
    if (t >= 43) {
      double acc = 0.0;
#pragma unroll
      for (int v = 0; v < 6; ++v) acc += t0c[v] * sg.CT[v][chQ];
      sg.Ch[chP][chQ] = acc;
      sg.Ch[chQ][chP] = acc;
    }
    if (t >= 37 && t < 43) {
      const int w = t - 37;
      double acc = 0.0;
#pragma unroll
      for (int v = 0; v < 6; ++v) acc += t0c[v] * sk.S[g][v];
      sg.TS[w] = acc;
    }
    __syncthreads();
    // D[i][c] += f_i f_c Ch[kRow_i][kRow_c]; L[i][col] += f_i Z[kRow_i][col]; Rt[i] += f_i TS[kRow_i];
    // R_u[3b+l] += (F S dV g_b)_l
    const double* fg = sg.f[g];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (t + 64 * j < 231) dacc[j] += fg[dI[j]] * fg[dC[j]] * sg.Ch[dKi[j]][dKc[j]];
    if (t < 48) {
#pragma unroll
      for (int m = 0; m < 11; ++m) {
        if (m < 10 || lh == 0) {
          const int kr = lh ? kRowOf(2 * m + 1) : kRowOf(2 * m);
          lacc[m] += fg[lh + 2 * m] * sg.Z[kr][lc];
        }
      }
    }
    if (t < NE) racc += fg[t] * sg.TS[rK];
  }
  if (t >= 40) {
    const int col = t - 40;
#pragma unroll
    for (int g = 0; g < 8; ++g) racc += sg.ru[g][col / 3][col % 3];
  }
  __syncthreads();  // the point data is dead from here: the condensation operands take its place
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (t + 64 * j < 231) {
      sc.A[dI[j]][dC[j]] = dacc[j];
      sc.A[dC[j]][dI[j]] = dacc[j];
    }
  }
  if (t < 48) {
#pragma unroll
    for (int m = 0; m < 11; ++m)
      if (m < 10 || lh == 0) sc.Lm[lh + 2 * m][lc] = lacc[m];
  }
  if (t < NE) sc.rt[t] = racc;
  if (t >= 40) sc.Ru[t - 40] = racc;
  // (2b) K_uu, lane (ka, kb) per 3x3 block (the K_uu inputs are untouched by the spill above)
  if constexpr (WANT_K) {
    double kb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll 1  // unrolled, the hoisted loads take registers (2 or 4: 166 / 154 VGPRs, no faster)
    for (int g = 0; g < 8; ++g) {
      double pa[3], pb[3], ga[3], gb[3], Ci6[6], S6[6], Ph[6];
      for (int i = 0; i < 3; ++i) {
        pa[i] = sk.P[g][ka][i];
        pb[i] = sk.P[g][kbn][i];
        ga[i] = sk.G[g][ka][i];
        gb[i] = sk.G[g][kbn][i];
      }
      for (int v = 0; v < 6; ++v) {
        Ci6[v] = sk.Ci[g][v];
        S6[v]  = sk.S[g][v];
        Ph[v]  = sk.Phi[g][v];
      }
      const double wl = sk.w[g][0], wc = sk.w[g][1];
      double qb[3], sb[3];
      symMul(Ci6, gb, qb);
      symMul(S6, gb, sb);
      const double gCg = ga[0] * qb[0] + ga[1] * qb[1] + ga[2] * qb[2];
      const double gSg = ga[0] * sb[0] + ga[1] * sb[1] + ga[2] * sb[2];
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l)
          kb[3 * k + l] += wl * pa[k] * pb[l] + wc * (pa[l] * pb[k] + gCg * Ph[v6(k, l)]) + (k == l ? gSg : 0.0);
    }
    __syncthreads();  // the K_uu inputs are dead: park the blocks in their place
    if (ka <= kbn)
#pragma unroll
      for (int i = 0; i < 9; ++i) sq.Kuu[upperBlock(ka, kbn)][i] = kb[i];
  }
  __syncthreads();

  // (3) The condensation L^T D^-1 [L | Rt] (enhancedassumedstrains.hh inverts D by LU, :243,294,344) through
  // D = L_D diag(d) L_D^T without pivoting (D is SPD while the material tangent is positive definite):
  // L^T D^-1 L = W^T diag(d)^-1 W with W = L_D^-1 L, so forward elimination of the augmented 21 x 46 matrix
  // [D | L | Rt] suffices. Lane c < 46 keeps column c in registers; step j reads rows j..20 of the pivot
  // column from lane j's registers (readlane), and every lane divides its row j, final from then on, by d_j and
  // subtracts the multiples from its rows below j. The L lanes first save their row j of W over the same row
  // of L in LDS. Afterwards the L lanes hold What = diag(d)^-1 W, the Rt lane what = diag(d)^-1 L_D^-1 Rt, and
  // lane c < 21 holds the multipliers (L_D)_cj = U_jc / d_j in rows j < c (U = diag(d) L_D^T, the eliminated
  // D). Half the lane reads and updates of a Gauss-Jordan elimination; the back substitution D^-1 needs for
  // the alpha update runs there, on 21 values per element.
  double col[NE];
#pragma unroll
  for (int i = 0; i < NE; ++i)
    col[i] = t < NE ? sc.A[i][t] : (t < NE + 24 ? sc.Lm[i][t - NE] : (t == NE + 24 ? sc.rt[i] : 0.0));
  // pivot guard: the reference inverts D with partial pivoting; without pivoting a pivot that is not finite or
  // vanishes against the largest entry of D makes the element singular here (error flag word 1, lowest element)
  double dmax = 0.0;
  if (t < NE)
#pragma unroll
    for (int i = 0; i < NE; ++i) dmax = fmax(dmax, fabs(col[i]));
  dmax = waveMax(dmax);
  bool singular = !(dmax < INFINITY);
#pragma unroll
  for (int j = 0; j < NE; ++j) {
    double piv[NE];  // rows j.. of lane j's column, read into scalar registers (no LDS round trip, no barrier)
#pragma unroll
    for (int i = j; i < NE; ++i) piv[i] = readLane(col[i], j);
    singular |= !(fabs(piv[j]) > 1e-13 * dmax && fabs(piv[j]) < INFINITY);
    if (t >= NE && t < NE + 24) sc.Lm[j][t - NE] = col[j];  // W_jc; lane c alone reads and writes column c
    const double f = col[j] / piv[j];
#pragma unroll
    for (int i = j + 1; i < NE; ++i) col[i] -= piv[i] * f;
    col[j] = f;
  }
  if (singular && t == 0) atomicMin(err + 1, (unsigned long long)e);
  // cache for the alpha update, per element: What column by column (entry (i, c) at 21 c + i, so that the
  // update's lane per mode reads 21 adjacent values at a time), what, then (L_D)_ij for i > j row by row
  // (offset 525 + i (i - 1) / 2 + j)
  double* cy = cache + e * kEasCache;
  if (t >= NE && t < NE + 24) {
#pragma unroll
    for (int i = 0; i < NE; ++i) {
      cy[(t - NE) * NE + i] = col[i];
      sc.A[i][t - NE]       = col[i];
    }
  } else if (t == NE + 24) {
#pragma unroll
    for (int i = 0; i < NE; ++i) {
      cy[NE * 24 + i] = col[i];
      sc.rt[i]        = col[i];
    }
  } else if (t > 0 && t < NE) {
    double* lrow = cy + NE * 25 + t * (t - 1) / 2;
#pragma unroll
    for (int j = 0; j < NE - 1; ++j)
      if (j < t) lrow[j] = col[j];
  }
  __syncthreads();
  // (4) K = K_uu - W^T What = K_uu - L^T D^-1 L, block (ka, kb) per lane; upper blocks stored, the diagonal
  // ones mirrored from their upper triangle (enhancedassumedstrains.hh:291-296)
  if constexpr (WANT_K) {
    const int blk = upperBlock(ka <= kbn ? ka : kbn, ka <= kbn ? kbn : ka);
    double kb[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) kb[i] = sq.Kuu[blk][i];
#pragma unroll
    for (int i = 0; i < NE; ++i) {
      double la[3], yb[3];
      for (int k = 0; k < 3; ++k) {
        la[k] = sc.Lm[i][3 * ka + k];
        yb[k] = sc.A[i][3 * kbn + k];
      }
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l) kb[3 * k + l] -= la[k] * yb[l];
    }
    if (ka == kbn) {
      kb[3] = kb[1];
      kb[6] = kb[2];
      kb[7] = kb[5];
    }
    if (ka <= kbn) {
      if (dest) {
        const int32_t code = dest[e * 36 + blk];
        double* dst        = Kebe + (int64_t)(code >> 1) * 9;
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) dst[3 * k + l] = (code & 1) ? kb[3 * l + k] : kb[3 * k + l];
      } else {
        double* dst = Kebe + (e * 36 + blk) * 9;
        for (int i = 0; i < 9; ++i) dst[i] = kb[i];
      }
    }
  }
  if (t >= 40) {
    const int r = t - 40;
    double acc  = sc.Ru[r];
#pragma unroll
    for (int i = 0; i < NE; ++i) acc -= sc.Lm[i][r] * sc.rt[i];
    const unsigned bits = fbits ? fbits[e] : 0u;
    if (bits) {  // Traction skill, after the EAS element's condensed internal forces
      const int k = r % 3;
      acc += faceLoad(X, conn, e, bits, r / 3, k == 0 ? trac.x : (k == 1 ? trac.y : trac.z));
    }
    Rebe[rdest ? (int64_t)rdest[e * 8 + r / 3] * 3 + r % 3 : e * 24 + r] = acc;
  }
}

// alpha_e -= D_e^-1 (Rt_e + L_e du_e) = L_D^-T (what_e + What_e du_e): 21 lanes per element, 3 elements per
// wavefront (lane 63 idle). Lane i forms z_i = what_i + (What du)_i; the back substitution with the unit upper
// triangle L_D^T then runs from the last mode up, x_k = z_k reaching lanes i < k of its element by a shuffle.
constexpr int kUpdElemsPerWave = 64 / NE;
__global__ __launch_bounds__(256) void easUpdateKernel(const double* __restrict__ cache,
                                                       const int32_t* __restrict__ conn,
                                                       const double* __restrict__ du, int64_t ne,
                                                       double* __restrict__ alpha) {
  const int lane = threadIdx.x & 63, sub = lane / NE, i = lane - NE * sub;
  const int64_t e   = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * kUpdElemsPerWave + sub;
  const bool active = sub < kUpdElemsPerWave && e < ne;
  const double* cy  = cache + (active ? e : 0) * kEasCache;
  double z          = 0.0;
  if (active) {
    z = cy[NE * 24 + i];
    for (int a = 0; a < 8; ++a) {
      const int64_t nd = conn[e * 8 + a];
      for (int k = 0; k < 3; ++k) z += cy[(a * 3 + k) * NE + i] * du[nd * 3 + k];
    }
  }
  // every lane takes part in the shuffles; lanes of other elements and idle lanes read values they drop
  for (int k = NE - 1; k > 0; --k) {
    const double xk = __shfl(z, NE * sub + k, 64);
    if (active && i < k) z -= cy[NE * 25 + k * (k - 1) / 2 + i] * xk;
  }
  if (active) alpha[e * NE + i] -= z;
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
  h->easCache.ensure((size_t)h->ne * kEasCache);
  if (wantK)
    launchEas<true>(h, e0, count, dest, rdest, h->ebeK64.p);
  else
    launchEas<false>(h, e0, count, nullptr, rdest, nullptr);
}

void easUpdate(ik_handle* h, const double* duDev) {
  require(h->easCacheValid, IK_EINVAL, "EAS update needs D, L, Rt of the current state (assemble first)");
  const int64_t perBlock = 4 * kUpdElemsPerWave;
  hipLaunchKernelGGL(easUpdateKernel, dim3((unsigned)((h->ne + perBlock - 1) / perBlock)), dim3(256), 0, h->stream, h->easCache.p,
                     h->conn.p, duDev, h->ne, h->alpha.p);
  h->easCacheValid = false;
}

}  // namespace ik
