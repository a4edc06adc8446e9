// assemble.hip — hex8 total-Lagrangian element residual/tangent and scatter into the block CSR matrix.
//
// Restates NonLinearElastic::calculateMatrixImpl / calculateVectorImpl
// (ikarus/finiteelements/mechanics/nonlinearelastic.hh:375-430) with the Neo-Hooke
// (materials/hyperelastic/neohooke.hh:79-142) and St.Venant-Kirchhoff (materials/svk.hh:78-164) laws,
// the Traction skill (mechanics/loads/traction.hh:106-138) and the SparseFlatAssembler scatter
// (ikarus/assembler/simpleassemblers.inl:121-169).
//
// Work decomposition (one 64-lane wavefront = 8 elements, 8 lanes per element):
//   phase 0  lane (el, q) stages node q of element el (X, u, Dirichlet flags) in LDS;
//   phase 1  lane (el, q) evaluates Gauss point q: J, J^-1, dN/dX for all 8 nodes, F, S*dV, C*dV (LDS);
//   phase 2  lane (el, a) accumulates block row a of K_e (8 blocks of 3x3) and R_e[a] over the 8 Gauss
//            points, reusing C*B_a for all b (the tangent is evaluated once per Gauss point, not per
//            node pair as at nonlinearelastic.hh:318);
//   phase 3  lane (el, a) adds its blocks into the block CSR values through the precomputed
//            block-slot map with hardware FP64/FP32 atomics (no sort, no coloring).
// The 8 Gauss points of a hex8 and its 8 nodes map onto the same 8 lanes, so no lane idles.
#include "ik_internal.hh"

namespace ik {
namespace {

constexpr int EPB = 8;  // elements per 64-lane block

// packed upper-triangle index of a symmetric 6x6 (row-major)
__host__ __device__ constexpr int sym6(int p, int q) {
  return p <= q ? p * 6 - (p * (p - 1)) / 2 + (q - p) : q * 6 - (q * (q - 1)) / 2 + (p - q);
}
// Voigt pairs [11,22,33,23,13,12] (ikarus/utils/tensorutils.hh:181-202)
__host__ __device__ constexpr int vI(int v) { return v < 3 ? v : (v == 3 ? 1 : 0); }
__host__ __device__ constexpr int vJ(int v) { return v < 3 ? v : (v == 5 ? 1 : 2); }

template <typename T>
__device__ inline void gaussPoint(int q, T xi[3]) {
  // 2-point Gauss-Legendre on [0,1]: 1/2 -+ 1/(2 sqrt 3), x-index fastest
  const T d = T(0.28867513459481288225);
#pragma unroll
  for (int k = 0; k < 3; ++k) xi[k] = ((q >> k) & 1) ? T(0.5) + d : T(0.5) - d;
}

// dN_a/dxi for the trilinear corner a = (a&1, a>>1&1, a>>2&1) (Dune reference cube)
template <typename T>
__device__ inline void dShape(int a, const T xi[3], T dN[3]) {
  T f[3], s[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const bool hi = (a >> k) & 1;
    f[k]          = hi ? xi[k] : T(1) - xi[k];
    s[k]          = hi ? T(1) : T(-1);
  }
  dN[0] = s[0] * f[1] * f[2];
  dN[1] = f[0] * s[1] * f[2];
  dN[2] = f[0] * f[1] * s[2];
}

template <typename T>
__device__ inline T det3(const T A[3][3]) {
  return A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
         A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
}

template <typename T>
__device__ inline void inv3(const T A[3][3], T d, T I[3][3]) {
  const T id = T(1) / d;
  I[0][0]    = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) * id;
  I[0][1]    = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
  I[0][2]    = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
  I[1][0]    = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) * id;
  I[1][1]    = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
  I[1][2]    = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
  I[2][0]    = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) * id;
  I[2][1]    = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
  I[2][2]    = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
}

// PK2 stress S (Voigt, no shear factor) and packed Voigt tangent from F.
// Returns false when det C <= 0 (the reference aborts, materialhelpers.hh:120-126).
template <typename T, int KIND>
__device__ inline bool material(T lam, T mu, const T F[3][3], T S[6], T Cp[21]) {
  T C[3][3];
#pragma unroll
  for (int I = 0; I < 3; ++I)
#pragma unroll
    for (int J = 0; J < 3; ++J) C[I][J] = F[0][I] * F[0][J] + F[1][I] * F[1][J] + F[2][I] * F[2][J];
  if constexpr (KIND == IK_SVK) {
    // E = (C - I)/2; S = lam tr(E) I + 2 mu E (svk.hh:100-140)
    const T E0 = T(0.5) * (C[0][0] - T(1)), E1 = T(0.5) * (C[1][1] - T(1)), E2 = T(0.5) * (C[2][2] - T(1));
    const T tr = E0 + E1 + E2;
    S[0]       = lam * tr + T(2) * mu * E0;
    S[1]       = lam * tr + T(2) * mu * E1;
    S[2]       = lam * tr + T(2) * mu * E2;
    S[3]       = mu * C[1][2];  // 2 mu E12 = mu C12
    S[4]       = mu * C[0][2];
    S[5]       = mu * C[0][1];
#pragma unroll
    for (int i = 0; i < 21; ++i) Cp[i] = T(0);
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = p; q < 3; ++q) Cp[sym6(p, q)] = lam + (p == q ? T(2) * mu : T(0));
#pragma unroll
    for (int p = 3; p < 6; ++p) Cp[sym6(p, p)] = mu;
    return true;
  } else {
    // Neo-Hooke (neohooke.hh:79-142): S = mu (I - C^-1) + lam lnJ C^-1,
    // C_IJKL = lam Ci_IJ Ci_KL + (mu - lam lnJ)(Ci_IK Ci_JL + Ci_IL Ci_JK)
    const T detC = det3(C);
    if (!(detC > T(0))) return false;
    T Ci[3][3];
    inv3(C, detC, Ci);
    const T lnJ = log(sqrt(detC));
    const T c2  = mu - lam * lnJ;
#pragma unroll
    for (int v = 0; v < 6; ++v) {
      const T ci = Ci[vI(v)][vJ(v)];
      S[v]       = mu * ((v < 3 ? T(1) : T(0)) - ci) + lam * lnJ * ci;
    }
#pragma unroll
    for (int p = 0; p < 6; ++p)
#pragma unroll
      for (int q = p; q < 6; ++q) {
        const int I = vI(p), J = vJ(p), K = vI(q), L = vJ(q);
        Cp[sym6(p, q)] = lam * Ci[I][J] * Ci[K][L] + c2 * (Ci[I][K] * Ci[J][L] + Ci[I][L] * Ci[J][K]);
      }
    return true;
  }
}

// B_a (6x3): column k = [F_k0 g0, F_k1 g1, F_k2 g2, F_k1 g2 + F_k2 g1, F_k0 g2 + F_k2 g0, F_k0 g1 + F_k1 g0]
// (derivative of the Green-Lagrange strain w.r.t. the nodal displacement, Voigt with engineering shear)
template <typename T>
__device__ inline void bop(const T F[3][3], const T g[3], T B[6][3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    B[0][k] = F[k][0] * g[0];
    B[1][k] = F[k][1] * g[1];
    B[2][k] = F[k][2] * g[2];
    B[3][k] = F[k][1] * g[2] + F[k][2] * g[1];
    B[4][k] = F[k][0] * g[2] + F[k][2] * g[0];
    B[5][k] = F[k][0] * g[1] + F[k][1] * g[0];
  }
}

template <typename T>
struct Lds {
  T X[EPB][8][3];
  T U[EPB][8][3];
  T G[EPB][8][8][3];  // [element][gauss point][node][dir]
  T F[EPB][8][9];
  T S[EPB][8][6];     // S * dV
  T C[EPB][8][21];    // C * dV (packed)
  int32_t node[EPB][8];
  uint8_t fix[EPB][24];
};

// One launch covers elements [e0, e0 + count). Kdense != nullptr writes K_e/R_e of element e0 densely
// (row-major 24x24, local order 3a+k) instead of scattering.
template <typename T, int KIND, bool WANT_K, bool WANT_R>
__global__ __launch_bounds__(64) void hex8Assemble(const double* __restrict__ X, const double* __restrict__ U,
                                                   const int32_t* __restrict__ conn,
                                                   const int32_t* __restrict__ bslot,
                                                   const uint8_t* __restrict__ fixed, int64_t e0, int64_t count,
                                                   int maskDbc, double lam_, double mu_, T* __restrict__ K,
                                                   double* __restrict__ R, unsigned long long* __restrict__ err,
                                                   double* __restrict__ Kdense, double* __restrict__ Rdense) {
  __shared__ Lds<T> s;
  const int lane     = threadIdx.x;
  const int el       = lane >> 3;
  const int q        = lane & 7;
  const int64_t eloc = (int64_t)blockIdx.x * EPB + el;
  const bool valid   = eloc < count;
  const int64_t e    = e0 + eloc;
  const T lam = T(lam_), mu = T(mu_);

  // phase 0: stage node q
  if (valid) {
    const int32_t nd = conn[e * 8 + q];
    s.node[el][q]    = nd;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      s.X[el][q][k]       = T(X[(int64_t)nd * 3 + k]);
      s.U[el][q][k]       = T(U[(int64_t)nd * 3 + k]);
      s.fix[el][q * 3 + k] = fixed[(int64_t)nd * 3 + k];
    }
  }
  __syncthreads();

  // phase 1: Gauss point q
  if (valid) {
    T xi[3];
    gaussPoint(q, xi);
    T dN[8][3];
#pragma unroll
    for (int a = 0; a < 8; ++a) dShape(a, xi, dN[a]);
    T J[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        T acc = T(0);
#pragma unroll
        for (int a = 0; a < 8; ++a) acc += s.X[el][a][i] * dN[a][j];
        J[i][j] = acc;
      }
    const T detJ = det3(J);
    T Ji[3][3];
    inv3(J, detJ, Ji);
    T F[3][3] = {{T(1), T(0), T(0)}, {T(0), T(1), T(0)}, {T(0), T(0), T(1)}};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      T g[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) g[i] = Ji[0][i] * dN[a][0] + Ji[1][i] * dN[a][1] + Ji[2][i] * dN[a][2];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        s.G[el][q][a][i] = g[i];
#pragma unroll
        for (int j = 0; j < 3; ++j) F[i][j] += s.U[el][a][i] * g[j];
      }
    }
    T Sv[6], Cp[21];
    const bool ok = material<T, KIND>(lam, mu, F, Sv, Cp);
    if (!ok) atomicMin(err, (unsigned long long)e);
    const T dV = fabs(detJ) * T(0.125);
#pragma unroll
    for (int i = 0; i < 9; ++i) s.F[el][q][i] = F[i / 3][i % 3];
#pragma unroll
    for (int v = 0; v < 6; ++v) s.S[el][q][v] = Sv[v] * dV;
#pragma unroll
    for (int i = 0; i < 21; ++i) s.C[el][q][i] = Cp[i] * dV;
  }
  __syncthreads();
  if (!valid) return;

  // phase 2: block row a = q
  const int a = q;
  T Kab[8][9];
  T Ra[3] = {T(0), T(0), T(0)};
  if constexpr (WANT_K) {
#pragma unroll
    for (int b = 0; b < 8; ++b)
#pragma unroll
      for (int i = 0; i < 9; ++i) Kab[b][i] = T(0);
  }
  for (int g = 0; g < 8; ++g) {
    T F[3][3], Sv[6], ga[3];
#pragma unroll
    for (int i = 0; i < 9; ++i) F[i / 3][i % 3] = s.F[el][g][i];
#pragma unroll
    for (int v = 0; v < 6; ++v) Sv[v] = s.S[el][g][v];
#pragma unroll
    for (int i = 0; i < 3; ++i) ga[i] = s.G[el][g][a][i];
    T Ba[6][3];
    bop(F, ga, Ba);
    if constexpr (WANT_R) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        T acc = T(0);
#pragma unroll
        for (int v = 0; v < 6; ++v) acc += Ba[v][k] * Sv[v];
        Ra[k] += acc;
      }
    }
    if constexpr (WANT_K) {
      T Cp[21];
#pragma unroll
      for (int i = 0; i < 21; ++i) Cp[i] = s.C[el][g][i];
      T CBa[6][3];
#pragma unroll
      for (int v = 0; v < 6; ++v)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          T acc = T(0);
#pragma unroll
          for (int w = 0; w < 6; ++w) acc += Cp[sym6(v, w)] * Ba[w][k];
          CBa[v][k] = acc;
        }
      // S g_a for the geometric stiffness g_a^T S g_b
      const T Sga0 = Sv[0] * ga[0] + Sv[5] * ga[1] + Sv[4] * ga[2];
      const T Sga1 = Sv[5] * ga[0] + Sv[1] * ga[1] + Sv[3] * ga[2];
      const T Sga2 = Sv[4] * ga[0] + Sv[3] * ga[1] + Sv[2] * ga[2];
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        T gb[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) gb[i] = s.G[el][g][b][i];
        T Bb[6][3];
        bop(F, gb, Bb);
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int l = 0; l < 3; ++l) {
            T acc = Kab[b][k * 3 + l];
#pragma unroll
            for (int v = 0; v < 6; ++v) acc += CBa[v][k] * Bb[v][l];
            Kab[b][k * 3 + l] = acc;
          }
        const T kg = Sga0 * gb[0] + Sga1 * gb[1] + Sga2 * gb[2];
        Kab[b][0] += kg;
        Kab[b][4] += kg;
        Kab[b][8] += kg;
      }
    }
  }

  // phase 3: scatter (or dense output for one element)
  if (Kdense) {
    if constexpr (WANT_K) {
#pragma unroll
      for (int b = 0; b < 8; ++b)
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int l = 0; l < 3; ++l) Kdense[(a * 3 + k) * 24 + b * 3 + l] = double(Kab[b][k * 3 + l]);
    }
    if constexpr (WANT_R) {
#pragma unroll
      for (int k = 0; k < 3; ++k) Rdense[a * 3 + k] = double(Ra[k]);
    }
    return;
  }
  if constexpr (WANT_K) {
    const int32_t* sl = bslot + e * 64 + a * 8;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      T* dst = K + (int64_t)sl[b] * 9;
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l) {
          if (maskDbc && (s.fix[el][a * 3 + k] | s.fix[el][b * 3 + l])) continue;
          unsafeAtomicAdd(dst + k * 3 + l, Kab[b][k * 3 + l]);
        }
    }
  }
  if constexpr (WANT_R) {
    const int64_t na = s.node[el][a];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (maskDbc && s.fix[el][a * 3 + k]) continue;
      unsafeAtomicAdd(R + na * 3 + k, double(Ra[k]));
    }
  }
}

// Traction faces (traction.hh:106-138): 1-point rule at the face centre, R_a -= N_a t |J_face|.
__global__ void tractionKernel(const double* __restrict__ X, const int32_t* __restrict__ conn,
                               const int32_t* __restrict__ faceElem, const int8_t* __restrict__ faceAxis,
                               const int8_t* __restrict__ faceSide, int64_t nf, double t0, double t1, double t2,
                               const uint8_t* __restrict__ fixed, int maskDbc, double* __restrict__ R) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int64_t e = faceElem[f];
  const int ax = faceAxis[f], sd = faceSide[f];
  double xi[3] = {0.5, 0.5, 0.5};
  xi[ax]       = (double)sd;
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  int32_t nodes[8];
  for (int a = 0; a < 8; ++a) {
    nodes[a] = conn[e * 8 + a];
    double dN[3];
    dShape(a, xi, dN);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) J[i][j] += X[(int64_t)nodes[a] * 3 + i] * dN[j];
  }
  const int ta = ax == 0 ? 1 : 0, tb = ax == 2 ? 1 : 2;
  const double c0 = J[1][ta] * J[2][tb] - J[2][ta] * J[1][tb];
  const double c1 = J[2][ta] * J[0][tb] - J[0][ta] * J[2][tb];
  const double c2 = J[0][ta] * J[1][tb] - J[1][ta] * J[0][tb];
  const double meas = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  const double t[3] = {t0, t1, t2};
  for (int a = 0; a < 8; ++a) {
    double N = 1.0;
    for (int k = 0; k < 3; ++k) N *= ((a >> k) & 1) ? xi[k] : 1.0 - xi[k];
    if (N == 0.0) continue;
    for (int k = 0; k < 3; ++k) {
      const int64_t dof = (int64_t)nodes[a] * 3 + k;
      if (maskDbc && fixed[dof]) continue;
      unsafeAtomicAdd(R + dof, -N * t[k] * meas);
    }
  }
}

__global__ void pointLoadKernel(const int64_t* __restrict__ dof, const double* __restrict__ coef, int64_t n,
                                double lambda, double* __restrict__ R) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) unsafeAtomicAdd(R + dof[i], coef[i] * lambda);
}

template <typename T>
__global__ void dbcDiagKernel(const int32_t* __restrict__ fixedList, int64_t n, const int32_t* __restrict__ diag,
                              T* __restrict__ K) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t dof = fixedList[i];
  const int32_t node = dof / 3, k = dof % 3;
  K[(int64_t)diag[node] * 9 + k * 4] = T(1);
}

template <typename T>
__global__ void fillKernel(T* __restrict__ p, int64_t n, T v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

inline unsigned gridFor(int64_t n, int bs) {
  int64_t g = (n + bs - 1) / bs;
  return (unsigned)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

template <typename T, int KIND, bool WK, bool WR>
void launchElem(ik_handle* h, int64_t e0, int64_t count, int mask, T* K, double* R, double* Kd, double* Rd) {
  const unsigned grid = (unsigned)((count + EPB - 1) / EPB);
  hipLaunchKernelGGL((hex8Assemble<T, KIND, WK, WR>), dim3(grid), dim3(64), 0, h->stream, h->X.p, h->u.p, h->conn.p,
                     h->bslot.p, h->fixed.p, e0, count, mask, h->lam, h->mu, K, R, h->errFlag.p, Kd, Rd);
}

template <typename T, bool WK, bool WR>
void launchElemMat(ik_handle* h, int64_t e0, int64_t count, int mask, T* K, double* R, double* Kd, double* Rd) {
  if (h->material == IK_SVK)
    launchElem<T, IK_SVK, WK, WR>(h, e0, count, mask, K, R, Kd, Rd);
  else
    launchElem<T, IK_NEOHOOKE, WK, WR>(h, e0, count, mask, K, R, Kd, Rd);
}

}  // namespace

void launchZero(double* p, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(fillKernel<double>, dim3(gridFor(n, 256)), dim3(256), 0, s, p, n, 0.0);
}
void launchZeroF(float* p, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(fillKernel<float>, dim3(gridFor(n, 256)), dim3(256), 0, s, p, n, 0.0f);
}

static void checkErrFlag(ik_handle* h) {
  unsigned long long v = 0;
  IK_HIP(hipMemcpyAsync(&v, h->errFlag.p, sizeof(v), hipMemcpyDeviceToHost, h->stream));
  IK_HIP(hipStreamSynchronize(h->stream));
  if (v != ~0ull) {
    h->bad_elem = (int64_t)v;
    unsigned long long reset = ~0ull;
    IK_HIP(hipMemcpy(h->errFlag.p, &reset, sizeof(reset), hipMemcpyHostToDevice));
    throw Error{IK_ENEGDET, "Determinant of right Cauchy Green tensor C must be greater than zero (element " +
                                std::to_string(v) + ")"};
  }
}

void assembleKR(ik_handle* h, int dbc, int precision, bool wantK, bool wantR) {
  require(h->neas == 0, IK_ENOTIMPL, "EAS assembly is handled by the EAS kernels");
  const int mask = dbc == IK_DBC_RAW ? 0 : 1;
  if (wantK) {
    if (precision == IK_FP64) {
      h->K64.ensure((size_t)h->nnzb * 9);
      launchZero(h->K64.p, h->nnzb * 9, h->stream);
    } else {
      h->K32.ensure((size_t)h->nnzb * 9);
      launchZeroF(h->K32.p, h->nnzb * 9, h->stream);
    }
  }
  if (wantR) launchZero(h->R.p, h->ndof, h->stream);
  if (wantK && precision == IK_FP64) {
    ScopedKernelTimer t(h, wantR ? KASM_KR : KASM_KR);
    if (wantR)
      launchElemMat<double, true, true>(h, 0, h->ne, mask, h->K64.p, h->R.p, nullptr, nullptr);
    else
      launchElemMat<double, true, false>(h, 0, h->ne, mask, h->K64.p, nullptr, nullptr, nullptr);
  } else {
    if (wantK) {
      ScopedKernelTimer t(h, KASM_KR);
      launchElemMat<float, true, false>(h, 0, h->ne, mask, h->K32.p, nullptr, nullptr, nullptr);
    }
    if (wantR) {
      ScopedKernelTimer t(h, KASM_R);
      launchElemMat<double, false, true>(h, 0, h->ne, mask, (double*)nullptr, h->R.p, nullptr, nullptr);
    }
  }
  if (wantR && h->nfaces) {
    hipLaunchKernelGGL(tractionKernel, dim3(gridFor(h->nfaces, 128)), dim3(128), 0, h->stream, h->X.p, h->conn.p,
                       h->faceElem.p, h->faceAxis.p, h->faceSide.p, h->nfaces, h->trac[0] * h->lambda,
                       h->trac[1] * h->lambda, h->trac[2] * h->lambda, h->fixed.p, mask, h->R.p);
  }
  if (wantR && h->npl) {
    hipLaunchKernelGGL(pointLoadKernel, dim3(gridFor(h->npl, 128)), dim3(128), 0, h->stream, h->plDof.p,
                       h->plCoef.p, h->npl, h->lambda, h->R.p);
  }
  if (wantK && dbc == IK_DBC_FULL) launchDbcDiag(h, precision);
  checkErrFlag(h);
}

void launchDbcDiag(ik_handle* h, int precision) {
  if (!h->nfixed) return;
  if (precision == IK_FP64)
    hipLaunchKernelGGL(dbcDiagKernel<double>, dim3(gridFor(h->nfixed, 256)), dim3(256), 0, h->stream,
                       h->fixedList.p, h->nfixed, h->diagBlk.p, h->K64.p);
  else
    hipLaunchKernelGGL(dbcDiagKernel<float>, dim3(gridFor(h->nfixed, 256)), dim3(256), 0, h->stream,
                       h->fixedList.p, h->nfixed, h->diagBlk.p, h->K32.p);
}

void elementHost(ik_handle* h, int64_t e, double* Ke, double* Re) {
  require(e >= 0 && e < h->ne, IK_EINVAL, "element index out of range");
  DevBuf<double> kd, rd;
  kd.ensure(576);
  rd.ensure(24);
  launchElemMat<double, true, true>(h, e, 1, 0, (double*)nullptr, nullptr, kd.p, rd.p);
  checkErrFlag(h);
  if (Ke) IK_HIP(hipMemcpyAsync(Ke, kd.p, 576 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  if (Re) IK_HIP(hipMemcpyAsync(Re, rd.p, 24 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  IK_HIP(hipStreamSynchronize(h->stream));
}

}  // namespace ik
