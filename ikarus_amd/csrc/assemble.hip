// assemble.hip — hex8 total-Lagrangian element residual/tangent and their assembly into the block CSR.
//
// Restates NonLinearElastic::calculateMatrixImpl / calculateVectorImpl
// (ikarus/finiteelements/mechanics/nonlinearelastic.hh:375-430) with the Neo-Hooke
// (materials/hyperelastic/neohooke.hh:79-142) and St.Venant-Kirchhoff (materials/svk.hh:78-164) laws,
// the Traction skill (mechanics/loads/traction.hh:106-138) and the SparseFlatAssembler scatter with its
// Dirichlet treatment (ikarus/assembler/simpleassemblers.inl:60-169).
//
// Element kernel (one 64-lane wavefront = 8 elements, 8 lanes per element):
//   phase 0  lane (el, q) stages node q of element el (X, u) in LDS;
//   phase 1  lane (el, q) evaluates Gauss point q: J, J^-1, dN/dX of all 8 nodes, F, S*dV, C*dV (LDS);
//   phase 2  lane (el, a) accumulates 4-5 of the 36 upper 3x3 blocks of K_e (blocks (a, a+j mod 8),
//            j = 0..3, plus j = 4 on lanes 0-3, so every lane has the same work) and R_e[a] over the
//            8 Gauss points; C*B_a is formed once per Gauss point and reused for every block;
//   phase 3  the blocks and R_e are stored element-by-element with plain stores.
// Gather kernels then sum, for every 3x3 block of the CSR pattern, the element blocks listed for it
// (a fixed order, so the result is bitwise reproducible), apply the FULL Dirichlet treatment on the fly
// and write each block once. Elements are processed in chunks; each chunk's gather runs right behind
// it on the rows the chunk completes, so the staged blocks are read back while still cached.
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
};

// upper-block index of (a, b), a <= b: row-major over the upper triangle of the 8x8 node pairs
__host__ __device__ constexpr int upperBlock(int a, int b) { return 8 * a - (a * (a - 1)) / 2 + (b - a); }

// One launch covers elements [e0, e0 + count). Outputs per element e (absolute index):
//   Kebe[(e*36 + sb)*9 + 3k + l] = K_e[3a+k][3b+l] for the upper block sb = upperBlock(a, b)
//   Rebe[e*24 + 3a + k]          = R_e[3a+k]
template <typename T, int KIND, bool WANT_K, bool WANT_R>
__global__ __launch_bounds__(64) void hex8Element(const double* __restrict__ X, const double* __restrict__ U,
                                                  const int32_t* __restrict__ conn, int64_t e0, int64_t count,
                                                  double lam_, double mu_, T* __restrict__ Kebe,
                                                  double* __restrict__ Rebe, unsigned long long* __restrict__ err) {
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
    const int64_t nd = conn[e * 8 + q];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      s.X[el][q][k] = T(X[nd * 3 + k]);
      s.U[el][q][k] = T(U[nd * 3 + k]);
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
    if constexpr (WANT_K) {
#pragma unroll
      for (int i = 0; i < 21; ++i) s.C[el][q][i] = Cp[i] * dV;
    }
  }
  __syncthreads();
  if (!valid) return;

  // phase 2: node a = q owns upper blocks (a, (a+j) & 7), j = 0..3 (+ j = 4 when a < 4)
  const int a = q;
  constexpr int NB = 5;
  const int nb = a < 4 ? 5 : 4;
  T Kb[NB][9];
  T Ra[3] = {T(0), T(0), T(0)};
  if constexpr (WANT_K) {
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int i = 0; i < 9; ++i) Kb[j][i] = T(0);
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
      T CBa[6][3];
#pragma unroll
      for (int v = 0; v < 6; ++v)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          T acc = T(0);
#pragma unroll
          for (int w = 0; w < 6; ++w) acc += s.C[el][g][sym6(v, w)] * Ba[w][k];
          CBa[v][k] = acc;
        }
      // S g_a for the geometric stiffness g_a^T S g_b
      const T Sga0 = Sv[0] * ga[0] + Sv[5] * ga[1] + Sv[4] * ga[2];
      const T Sga1 = Sv[5] * ga[0] + Sv[1] * ga[1] + Sv[3] * ga[2];
      const T Sga2 = Sv[4] * ga[0] + Sv[3] * ga[1] + Sv[2] * ga[2];
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        if (j < nb) {
          const int b = (a + j) & 7;
          T gb[3];
#pragma unroll
          for (int i = 0; i < 3; ++i) gb[i] = s.G[el][g][b][i];
          T Bb[6][3];
          bop(F, gb, Bb);
#pragma unroll
          for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int l = 0; l < 3; ++l) {
              T acc = Kb[j][k * 3 + l];
#pragma unroll
              for (int v = 0; v < 6; ++v) acc += CBa[v][k] * Bb[v][l];
              Kb[j][k * 3 + l] = acc;
            }
          const T kg = Sga0 * gb[0] + Sga1 * gb[1] + Sga2 * gb[2];
          Kb[j][0] += kg;
          Kb[j][4] += kg;
          Kb[j][8] += kg;
        }
      }
    }
  }

  // phase 3: element-by-element stores (K_e(a,b) for b >= a, else transposed into (b,a))
  if constexpr (WANT_K) {
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < nb) {
        const int b    = (a + j) & 7;
        const bool tr  = b < a;
        const int sb   = tr ? upperBlock(b, a) : upperBlock(a, b);
        T* dst         = Kebe + (e * 36 + sb) * 9;
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int l = 0; l < 3; ++l) dst[tr ? l * 3 + k : k * 3 + l] = Kb[j][k * 3 + l];
      }
    }
  }
  if constexpr (WANT_R) {
#pragma unroll
    for (int k = 0; k < 3; ++k) Rebe[e * 24 + a * 3 + k] = double(Ra[k]);
  }
}

// K gather: thread per 3x3 block over blocks [p0, p1): sum the listed element blocks (transposed where
// flagged) in list order, then FULL: rows/columns of constrained DOFs are zero with a unit diagonal
// (simpleassemblers.inl:152-169).
template <typename T>
__global__ __launch_bounds__(256) void gatherK(const int32_t* __restrict__ kcptr, const int32_t* __restrict__ kclist,
                                               const int32_t* __restrict__ blockRow,
                                               const int32_t* __restrict__ colidx,
                                               const uint8_t* __restrict__ fixed, int maskDbc,
                                               const T* __restrict__ Kebe, int64_t p0, int64_t p1,
                                               T* __restrict__ K) {
  const int64_t p = p0 + (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= p1) return;
  const int32_t c0 = kcptr[p], c1 = kcptr[p + 1];
  T acc[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) acc[i] = T(0);
  for (int32_t c = c0; c < c1; ++c) {
    const int32_t code = kclist[c];
    const T* src       = Kebe + (int64_t)(code >> 1) * 9;
    T v[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) v[i] = src[i];
    const bool tr = code & 1;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int l = 0; l < 3; ++l) acc[k * 3 + l] += tr ? v[l * 3 + k] : v[k * 3 + l];
  }
  if (maskDbc) {
    const int64_t r = (int64_t)blockRow[p] * 3, c = (int64_t)colidx[p] * 3;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int l = 0; l < 3; ++l)
        if (fixed[r + k] | fixed[c + l]) acc[k * 3 + l] = (r + k == c + l) ? T(1) : T(0);
  }
  T* dst = K + p * 9;
#pragma unroll
  for (int i = 0; i < 9; ++i) dst[i] = acc[i];
}

// R gather: thread per DOF over node rows [r0, r1)
__global__ __launch_bounds__(256) void gatherR(const int32_t* __restrict__ ncptr, const int32_t* __restrict__ nclist,
                                               const uint8_t* __restrict__ fixed, int maskDbc,
                                               const double* __restrict__ Rebe, int64_t r0, int64_t r1,
                                               double* __restrict__ R) {
  const int64_t d = r0 * 3 + (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (d >= r1 * 3) return;
  if (maskDbc && fixed[d]) {
    R[d] = 0.0;
    return;
  }
  const int64_t node = d / 3;
  const int k        = (int)(d - node * 3);
  double acc         = 0.0;
  for (int32_t c = ncptr[node]; c < ncptr[node + 1]; ++c) {
    const int32_t code = nclist[c];  // e*8 + a
    acc += Rebe[(int64_t)(code >> 3) * 24 + (code & 7) * 3 + k];
  }
  R[d] = acc;
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
__global__ void fillKernel(T* __restrict__ p, int64_t n, T v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// dense 24x24 / 24 output of one element from its staged blocks (calculateMatrix/calculateVector)
__global__ void unpackElement(const double* __restrict__ Kebe, const double* __restrict__ Rebe, int64_t e,
                              double* __restrict__ Kd, double* __restrict__ Rd) {
  const int t = threadIdx.x;  // 576 threads
  if (t < 576) {
    const int r = t / 24, c = t % 24;
    const int a = r / 3, k = r % 3, b = c / 3, l = c % 3;
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    const double* blk = Kebe + (e * 36 + upperBlock(lo, hi)) * 9;
    Kd[t] = a <= b ? blk[k * 3 + l] : blk[l * 3 + k];
  }
  if (t < 24) Rd[t] = Rebe[e * 24 + t];
}

inline unsigned gridFor(int64_t n, int bs) {
  int64_t g = (n + bs - 1) / bs;
  return (unsigned)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

template <typename T, int KIND, bool WK, bool WR>
void launchElem(ik_handle* h, int64_t e0, int64_t count, T* Kebe, double* Rebe) {
  const unsigned grid = (unsigned)((count + EPB - 1) / EPB);
  hipLaunchKernelGGL((hex8Element<T, KIND, WK, WR>), dim3(grid), dim3(64), 0, h->stream, h->X.p, h->u.p, h->conn.p,
                     e0, count, h->lam, h->mu, Kebe, Rebe, h->errFlag.p);
}

template <typename T, bool WK, bool WR>
void launchElemMat(ik_handle* h, int64_t e0, int64_t count, T* Kebe, double* Rebe) {
  if (h->material == IK_SVK)
    launchElem<T, IK_SVK, WK, WR>(h, e0, count, Kebe, Rebe);
  else
    launchElem<T, IK_NEOHOOKE, WK, WR>(h, e0, count, Kebe, Rebe);
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

// Element chunks on the main stream; each chunk's gathers on the second stream, released by an event
// recorded behind the chunk, so gathers (memory-bound) overlap the next chunks (compute-bound).
template <typename T>
static void runChunks(ik_handle* h, int dbc, bool wantK, bool wantR, T* Kebe, T* K) {
  const int mask = dbc == IK_DBC_RAW ? 0 : 1;
  const size_t nc = h->chunkElem.size() - 1;
  while (h->chunkEvents.size() < nc) {
    hipEvent_t ev;
    IK_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    h->chunkEvents.push_back(ev);
  }
  for (size_t c = 0; c < nc; ++c) {
    const int64_t e0 = h->chunkElem[c], e1 = h->chunkElem[c + 1];
    {
      ScopedKernelTimer t(h, wantK ? KASM_KR : KASM_R, h->stream);
      if (h->neas)
        easStage(h, e0, e1 - e0, wantK);
      else if (wantK && wantR)
        launchElemMat<T, true, true>(h, e0, e1 - e0, Kebe, h->ebeR.p);
      else if (wantK)
        launchElemMat<T, true, false>(h, e0, e1 - e0, Kebe, nullptr);
      else
        launchElemMat<double, false, true>(h, e0, e1 - e0, (double*)nullptr, h->ebeR.p);
    }
    const int64_t r0 = h->chunkRow[c], r1 = h->chunkRow[c + 1];
    if (r1 <= r0) continue;
    IK_HIP(hipEventRecord(h->chunkEvents[c], h->stream));
    IK_HIP(hipStreamWaitEvent(h->stream2, h->chunkEvents[c], 0));
    if (wantK) {
      const int64_t p0 = h->hrowptr[r0], p1 = h->hrowptr[r1];
      ScopedKernelTimer t(h, KGATHER, h->stream2);
      hipLaunchKernelGGL(gatherK<T>, dim3((unsigned)((p1 - p0 + 255) / 256)), dim3(256), 0, h->stream2, h->kcptr.p,
                         h->kclist.p, h->blockRow.p, h->colidx.p, h->fixed.p, mask, Kebe, p0, p1, K);
    }
    if (wantR)
      hipLaunchKernelGGL(gatherR, dim3((unsigned)(((r1 - r0) * 3 + 255) / 256)), dim3(256), 0, h->stream2,
                         h->ncptr.p, h->nclist.p, h->fixed.p, mask, h->ebeR.p, r0, r1, h->R.p);
  }
  // join: later work on the main stream sees every gathered row
  IK_HIP(hipEventRecord(h->joinEvent, h->stream2));
  IK_HIP(hipStreamWaitEvent(h->stream, h->joinEvent, 0));
}

void assembleKR(ik_handle* h, int dbc, int precision, bool wantK, bool wantR) {
  const int mask = dbc == IK_DBC_RAW ? 0 : 1;
  h->ebeR.ensure((size_t)h->ne * 24);  // EAS elements always produce R_e
  if (h->neas) {
    require(precision == IK_FP64, IK_ENOTIMPL, "the EAS element is FP64 only");
    // the EAS kernel computes the condensed R_e with K_e; both gathers run when K is wanted
    wantR = true;
  }
  // the mixed form pairs an FP32 tangent with an FP64 residual: two passes
  const bool fused = wantK && wantR && precision == IK_FP64;
  if (wantK && precision == IK_FP64) {
    h->K64.ensure((size_t)h->nnzb * 9);
    h->ebeK64.ensure((size_t)h->ne * 324);
    runChunks<double>(h, dbc, true, fused, h->ebeK64.p, h->K64.p);
  } else if (wantK) {
    h->K32.ensure((size_t)h->nnzb * 9);
    h->ebeK32.ensure((size_t)h->ne * 324);
    runChunks<float>(h, dbc, true, false, h->ebeK32.p, h->K32.p);
  }
  if (wantR && !fused) runChunks<double>(h, dbc, false, true, (double*)nullptr, (double*)nullptr);
  if (h->neas) h->easCacheValid = true;
  if (wantR && h->nfaces) {
    hipLaunchKernelGGL(tractionKernel, dim3(gridFor(h->nfaces, 128)), dim3(128), 0, h->stream, h->X.p, h->conn.p,
                       h->faceElem.p, h->faceAxis.p, h->faceSide.p, h->nfaces, h->trac[0] * h->lambda,
                       h->trac[1] * h->lambda, h->trac[2] * h->lambda, h->fixed.p, mask, h->R.p);
  }
  if (wantR && h->npl) {
    hipLaunchKernelGGL(pointLoadKernel, dim3(gridFor(h->npl, 128)), dim3(128), 0, h->stream, h->plDof.p,
                       h->plCoef.p, h->npl, h->lambda, h->R.p);
  }
  checkErrFlag(h);
}

void elementHost(ik_handle* h, int64_t e, double* Ke, double* Re) {
  require(e >= 0 && e < h->ne, IK_EINVAL, "element index out of range");
  h->ebeK64.ensure((size_t)h->ne * 324);
  h->ebeR.ensure((size_t)h->ne * 24);
  if (h->neas)
    easStage(h, e, 1, true);
  else
    launchElemMat<double, true, true>(h, e, 1, h->ebeK64.p, h->ebeR.p);
  checkErrFlag(h);
  DevBuf<double> kd, rd;
  kd.ensure(576);
  rd.ensure(24);
  hipLaunchKernelGGL(unpackElement, dim3(1), dim3(576), 0, h->stream, h->ebeK64.p, h->ebeR.p, e, kd.p, rd.p);
  if (Ke) IK_HIP(hipMemcpyAsync(Ke, kd.p, 576 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  if (Re) IK_HIP(hipMemcpyAsync(Re, rd.p, 24 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  IK_HIP(hipStreamSynchronize(h->stream));
}

}  // namespace ik
