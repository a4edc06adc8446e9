// assemble.hip — hex8 total-Lagrangian element residual/tangent and their assembly into the block CSR.
//
// Restates NonLinearElastic::calculateMatrixImpl / calculateVectorImpl
// (ikarus/finiteelements/mechanics/nonlinearelastic.hh:375-430) with the Neo-Hooke
// (materials/hyperelastic/neohooke.hh:79-142) and St.Venant-Kirchhoff (materials/svk.hh:78-164) laws,
// the Traction skill (mechanics/loads/traction.hh:106-138) and the SparseFlatAssembler scatter with its
// Dirichlet treatment (ikarus/assembler/simpleassemblers.inl:60-169).
//
// Tangent form. With g_a = dN_a/dX, C^-1 = Ci and the per-node vectors p_a = F Ci g_a, the Voigt product
// B_a^T CC B_b + (g_a^T S g_b) I of both laws collapses (minor symmetries of CC) to
//   K_ab[k][l] = dV [ w1 p_ak p_bl + w2 p_al p_bk + w2 (g_a . Ci g_b) Phi_kl + (g_a . S g_b) d_kl ],
//   Phi = F Ci F^T,  Neo-Hooke: w1 = lambda, w2 = mu - lambda ln J;  SVK: Ci -> I, w1 = lambda, w2 = mu,
// which is the reference's B^T CC B + kg (nonlinearelastic.hh:313-322, 294-302) summed in another order.
// The two rank-1 parts are batched 24x24 contractions over the 8 Gauss points,
//   W1 = sum_g w1 dV p_g p_g^T,  W2 = sum_g w2 dV p_g p_g^T,  K_ab[k][l] = W1[3a+k][3b+l] + W2[3a+l][3b+k] + ...,
// and run on the FP64 matrix cores (v_mfma_f64_4x4x4: the four 4x4 blocks of one instruction are four
// elements; 21 upper tiles x 2 k-steps per matrix).
//
// Tangent kernel (one workgroup = a tile of 16 elements, 4 per 64-lane wavefront):
//   each lane evaluates two Gauss points of one element for two of its nodes (J and F partial sums
//   combined across the lane quad with DPP), then the per-point material; the tangent's terms are
//   accumulated on the FP64 matrix cores (operands as below); the element's upper 3x3 blocks go to LDS at
//   the position of the K block they fall on, and each K block's sum over the tile's elements is staged.
// Gather kernels then sum, for every 3x3 block of the CSR pattern, the staged blocks listed for it (a
// fixed order, so the result is bitwise reproducible), apply the FULL Dirichlet treatment on the fly and
// write each block once. Large meshes are processed in chunks whose gathers run behind them on the rows
// each chunk completes.
#include "ik_internal.hh"

#include <type_traits>

namespace ik {
namespace {

constexpr int EPW = 4;  // elements per 64-lane wavefront

// Voigt pairs [11,22,33,23,13,12] (ikarus/utils/tensorutils.hh:181-202)
__host__ __device__ constexpr int vI(int v) { return v < 3 ? v : (v == 3 ? 1 : 0); }
__host__ __device__ constexpr int vJ(int v) { return v < 3 ? v : (v == 5 ? 1 : 2); }
// Voigt index of the symmetric pair (i, j)
__host__ __device__ constexpr int voigt(int i, int j) { return i == j ? i : 6 - i - j; }

__device__ inline void gaussPoint(int q, double xi[3]) {
  // 2-point Gauss-Legendre on [0,1]: 1/2 -+ 1/(2 sqrt 3), x-index fastest
  const double d = 0.28867513459481288225;
#pragma unroll
  for (int k = 0; k < 3; ++k) xi[k] = ((q >> k) & 1) ? 0.5 + d : 0.5 - d;
}

// dN_a/dxi for the trilinear corner a = (a&1, a>>1&1, a>>2&1) (Dune reference cube)
__device__ inline void dShape(int a, const double xi[3], double dN[3]) {
  double f[3], s[3];
#pragma unroll
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

__device__ inline void inv3(const double A[3][3], double d, double I[3][3]) {
  const double id = 1.0 / d;
  I[0][0]         = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) * id;
  I[0][1]         = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
  I[0][2]         = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
  I[1][0]         = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) * id;
  I[1][1]         = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
  I[1][2]         = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
  I[2][0]         = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) * id;
  I[2][1]         = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
  I[2][2]         = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
}

// PK2 stress S (Voigt), Ci = C^-1 (Voigt-packed, Neo-Hooke only) and the weight w2 of the tangent form.
// Returns false when det C <= 0 (the reference aborts, materialhelpers.hh:120-126).
template <int KIND>
__device__ inline bool material(double lam, double mu, const double F[3][3], double S[6], double Ci[6],
                                double& w2) {
  double C[6];
#pragma unroll
  for (int v = 0; v < 6; ++v) C[v] = F[0][vI(v)] * F[0][vJ(v)] + F[1][vI(v)] * F[1][vJ(v)] + F[2][vI(v)] * F[2][vJ(v)];
  if constexpr (KIND == IK_SVK) {
    // E = (C - I)/2; S = lam tr(E) I + 2 mu E (svk.hh:100-140)
    const double E0 = 0.5 * (C[0] - 1.0), E1 = 0.5 * (C[1] - 1.0), E2 = 0.5 * (C[2] - 1.0);
    const double tr = E0 + E1 + E2;
    S[0]            = lam * tr + 2.0 * mu * E0;
    S[1]            = lam * tr + 2.0 * mu * E1;
    S[2]            = lam * tr + 2.0 * mu * E2;
    S[3]            = mu * C[3];  // 2 mu E23 = mu C23
    S[4]            = mu * C[4];
    S[5]            = mu * C[5];
#pragma unroll
    for (int v = 0; v < 6; ++v) Ci[v] = v < 3 ? 1.0 : 0.0;
    w2 = mu;
    return true;
  } else {
    // Neo-Hooke (neohooke.hh:79-142): S = mu (I - C^-1) + lam lnJ C^-1
    const double detC = C[0] * (C[1] * C[2] - C[3] * C[3]) - C[5] * (C[5] * C[2] - C[3] * C[4]) +
                        C[4] * (C[5] * C[3] - C[1] * C[4]);
    if (!(detC > 0.0)) return false;
    const double id = 1.0 / detC;
    Ci[0]           = (C[1] * C[2] - C[3] * C[3]) * id;
    Ci[1]           = (C[0] * C[2] - C[4] * C[4]) * id;
    Ci[2]           = (C[0] * C[1] - C[5] * C[5]) * id;
    Ci[3]           = (C[4] * C[5] - C[0] * C[3]) * id;
    Ci[4]           = (C[5] * C[3] - C[1] * C[4]) * id;
    Ci[5]           = (C[4] * C[3] - C[2] * C[5]) * id;
    const double lnJ = log(sqrt(detC));
#pragma unroll
    for (int v = 0; v < 6; ++v) S[v] = mu * ((v < 3 ? 1.0 : 0.0) - Ci[v]) + lam * lnJ * Ci[v];
    w2 = mu - lam * lnJ;
    return true;
  }
}

// symmetric (Voigt-packed) matrix times vector
__device__ inline void symMul(const double M[6], const double x[3], double y[3]) {
  y[0] = M[0] * x[0] + M[5] * x[1] + M[4] * x[2];
  y[1] = M[5] * x[0] + M[1] * x[1] + M[3] * x[2];
  y[2] = M[4] * x[0] + M[3] * x[1] + M[2] * x[2];
}

// DPP lane moves within a row of 16 lanes (64-bit value as two 32-bit moves)
template <int CTRL>
__device__ inline double dpp(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo);
}
constexpr int DPP_XOR1 = 0xB1;   // quad_perm [1,0,3,2]
constexpr int DPP_XOR2 = 0x4E;   // quad_perm [2,3,0,1]
constexpr int DPP_XOR3 = 0x1B;   // quad_perm [3,2,1,0]
constexpr int DPP_HALF_MIRROR = 0x141;  // row_half_mirror: lane q of each group of 8 reads lane 7 - q

// value of lane (lane ^ BIT), BIT = 1, 2 or 4, by DPP moves (q ^ 4 = (7 - q) ^ 3 within a group of 8)
template <int BIT>
__device__ inline double laneXor(double v) {
  if constexpr (BIT == 1) return dpp<DPP_XOR1>(v);
  else if constexpr (BIT == 2) return dpp<DPP_XOR2>(v);
  else return dpp<DPP_XOR3>(dpp<DPP_HALF_MIRROR>(v));
}

// upper-block index of (a, b), a <= b: row-major over the upper triangle of the 8x8 node pairs
__host__ __device__ constexpr int upperBlock(int a, int b) { return 8 * a - (a * (a - 1)) / 2 + (b - a); }

// copy n staged values (n a multiple of 4) to global memory with 16-byte stores
template <typename TK>
__device__ inline void storeBlocks(const double* src, TK* dst, int n, int lane) {
  if constexpr (sizeof(TK) == 8) {
    const double2* s = reinterpret_cast<const double2*>(src);
    double2* d       = reinterpret_cast<double2*>(dst);
    for (int j = lane; j < n / 2; j += 64) d[j] = s[j];
  } else {
    const double4* s = reinterpret_cast<const double4*>(src);
    float4* d        = reinterpret_cast<float4*>(dst);
    for (int j = lane; j < n / 4; j += 64) {
      const double4 v = s[j];
      d[j]            = make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
    }
  }
}

// symmetric-pair index of (k, l) = its Voigt index
constexpr int NSYM = 6;

// Workgroup = 4 wavefronts = one tile of kTileElems elements, EPW per wavefront.
//  TILED: the workgroup takes tile t = e0 + blockIdx.x, element s = tileElem[16 t + s] (-1: padding). Its
//   576 upper element blocks are summed in LDS per group (the blocks falling on one K block, in the K
//   block's orientation): rank 0 of every group stores, ranks 1, 2, ... add in turn, one rank per barrier
//   step, so each sum has a fixed order and no two lanes touch one value in a step. Each group sum then
//   goes to its staging slot gdst[g] or, for a K block no other tile touches, straight into K
//   (FULL-treated when maskDbc, and to its mirror block).
//  otherwise: elements e0 + 16 blockIdx.x + s below e0 + count, stored in element order,
//   Kebe[(e*36 + sb)*9 + 3k + l] = K_e[3a+k][3b+l] for the upper block sb = upperBlock(a, b).
// WANT_R: Rebe[rdest ? rdest[e*8+a]*3 + k : e*24 + 3a + k] = R_e[3a+k] (element order only).
// The element values are computed in FP64; TK = float rounds the group sums once at the store.
//
// Lane = 16 kk + 4 bk + i of wavefront w: element s = 4 w + bk, Gauss points kk and kk + 4, nodes i and
// i + 4. In the component-major ordering (row 8k + a) of K_e, the 8x8 node matrix K^{kl}[a][b] =
// K_ab[k][l] is
//   K^{kl} = sum_g [ w1 p_k p_l^T + w2 p_l p_k^T ] + sum_{g,I} g_I (Psi_kl Ci g + d_kl S g)_I^T,
// (p_k = (p_ak)_a, g_I = (g_aI)_a over the nodes) so with the contraction index (gauss point, term) every
// operand of v_mfma_f64_4x4x4 is a value the lane computed itself: A[a][k] at lane 16 k + 4 bk + a,
// B[k][b] at lane 16 k + 4 bk + b, D[a][b] at lane 16 a + 4 bk + b. Tiles (alpha, beta) of K^{kl} cover
// nodes 4 alpha..+3 x 4 beta..+3; (0,1) for all kl, (0,0) and (1,1) for k <= l (the rest follow from
// K_ab[k][l] = K_ba[l][k]): 21 tiles x (2 gauss blocks x (2 W terms + 3 I)) contraction steps.
template <int KIND, typename TK, bool WANT_K, bool WANT_R, bool TILED>
__global__ __launch_bounds__(kTileThreads, 3) void hex8Element(const double* __restrict__ X, const double* __restrict__ U,
                                                      const int32_t* __restrict__ conn,
                                                      const int32_t* __restrict__ tileElem, int64_t e0,
                                                      int64_t count, double lam, double mu,
                                                      const int32_t* __restrict__ tgptr,
                                                      const int32_t* __restrict__ gdst,
                                                      const int32_t* __restrict__ gpt,
                                                      const uint16_t* __restrict__ gmsk,
                                                      const uint16_t* __restrict__ epos,
                                                      const uint8_t* __restrict__ tmaxr, TK* __restrict__ K,
                                                      int maskDbc,
                                                      const int32_t* __restrict__ rdest, TK* __restrict__ Kebe,
                                                      double* __restrict__ Rebe, const uint8_t* __restrict__ fbits,
                                                      double3 trac, unsigned long long* __restrict__ err) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kk = lane >> 4, bk = (lane >> 2) & 3, i = lane & 3;
  const int s  = EPW * wave + bk;  // element of the tile
  const int64_t t  = e0 + blockIdx.x;
  int64_t e;
  bool valid;
  int nvalidWave = 0;  // element-order stores: valid elements of this wavefront
  if constexpr (TILED) {
    const int32_t ev = tileElem[t * kTileElems + s];
    valid            = ev >= 0;
    // lanes of a padding slot compute the tile's first element; only their stores are dropped
    e = valid ? ev : tileElem[t * kTileElems];
  } else {
    const int64_t base = (int64_t)blockIdx.x * kTileElems + EPW * wave;
    nvalidWave         = (int)(count - base < EPW ? (count - base < 0 ? 0 : count - base) : EPW);
    valid              = bk < nvalidWave;
    const int64_t idx  = (int64_t)blockIdx.x * kTileElems + s;
    e                  = e0 + (idx < count ? idx : count - 1);
  }

  // the tile's entry groups and ranks and its group targets, fetched early so the stores do not wait on them
  __shared__ uint16_t ep[kTileEntries], gm[kTileEntries];
  __shared__ int32_t gs[kTileEntries], gt[kTileEntries];
  int ng = 0, nrank = 1;
  if constexpr (WANT_K && TILED) {
    const int32_t g0 = tgptr[t];
    ng               = tgptr[t + 1] - g0;
    nrank            = tmaxr[t];
    for (int j = threadIdx.x; j < kTileEntries; j += kTileThreads) ep[j] = epos[t * kTileEntries + j];
    for (int j = threadIdx.x; j < ng; j += kTileThreads) {
      gs[j] = gdst[g0 + j];
      gt[j] = gpt[g0 + j];
      gm[j] = gmsk[g0 + j];
    }
  }

  double x[2][3], u[2][3];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t nd = conn[e * 8 + i + 4 * h];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      x[h][k] = X[nd * 3 + k];
      u[h][k] = U[nd * 3 + k];
    }
  }

  // accumulators: K01[3k+l] tile (0,1) of K^{kl}; K00/K11[v] tiles (0,0)/(1,1) of K^{kl}, k <= l, v = voigt(k,l);
  // G01[v - 3] the Psi term of tile (0,1) for k < l (shared by K^{kl} and K^{lk})
  double K01[9], K00[NSYM], K11[NSYM], G01[3];
  // residual: Rq[h] at lane 16 a + 4 bk + c accumulates R_e[3 (a + 4 h) + c], c < 3
  double Rq[2] = {0.0, 0.0};
  if constexpr (WANT_K) {
#pragma unroll
    for (int t = 0; t < 9; ++t) K01[t] = 0.0;
#pragma unroll
    for (int t = 0; t < NSYM; ++t) K00[t] = K11[t] = 0.0;
#pragma unroll
    for (int t = 0; t < 3; ++t) G01[t] = 0.0;
  }

#pragma unroll
  for (int sblk = 0; sblk < 2; ++sblk) {
    const int gp = 4 * sblk + kk;
    double xi[3];
    gaussPoint(gp, xi);
    double dN[2][3];
#pragma unroll
    for (int h = 0; h < 2; ++h) dShape(i + 4 * h, xi, dN[h]);
    // J = sum_a x_a dN_a^T over the quad's 8 nodes (2 per lane, combined with DPP)
    double J[3][3];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        double v = x[0][p] * dN[0][q] + x[1][p] * dN[1][q];
        v += dpp<DPP_XOR1>(v);
        v += dpp<DPP_XOR2>(v);
        J[p][q] = v;
      }
    const double detJ = det3(J);
    double Ji[3][3];
    inv3(J, detJ, Ji);
    double g[2][3];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int p = 0; p < 3; ++p) g[h][p] = Ji[0][p] * dN[h][0] + Ji[1][p] * dN[h][1] + Ji[2][p] * dN[h][2];
    double F[3][3];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        double v = u[0][p] * g[0][q] + u[1][p] * g[1][q];
        v += dpp<DPP_XOR1>(v);
        v += dpp<DPP_XOR2>(v);
        F[p][q] = (p == q ? 1.0 : 0.0) + v;
      }
    double Sv[6], Ci[6], w2;
    const bool ok = material<KIND>(lam, mu, F, Sv, Ci, w2);
    if (!ok && valid && i == 0) atomicMin(err, (unsigned long long)e);
    const double dV = fabs(detJ) * 0.125;
#pragma unroll
    for (int v = 0; v < 6; ++v) Sv[v] *= dV;
    double sg[2][3];  // S dV g
#pragma unroll
    for (int h = 0; h < 2; ++h) symMul(Sv, g[h], sg[h]);
    if constexpr (WANT_R) {
      // R_ac = sum_gp sum_I g_aI T_cI with T = F S dV: A[a][gp] = g_aI (this lane's node, Gauss point),
      // B[gp][c] = T_cI (row c = i of T; zero for i = 3), accumulated over I and both Gauss-point blocks
      double Fr[3], T[3];
#pragma unroll
      for (int J = 0; J < 3; ++J) Fr[J] = i == 0 ? F[0][J] : (i == 1 ? F[1][J] : (i == 2 ? F[2][J] : 0.0));
      symMul(Sv, Fr, T);
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int I = 0; I < 3; ++I) Rq[h] = __builtin_amdgcn_mfma_f64_4x4x4f64(g[h][I], T[I], Rq[h], 0, 0, 0);
    }
    if constexpr (WANT_K) {
      // M = F Ci, p = M g, Psi = w2 dV M F^T, cg = Ci g
      double M[3][3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double row[3] = {F[k][0], F[k][1], F[k][2]};
        symMul(Ci, row, M[k]);
      }
      const double w1d = lam * dV, w2d = w2 * dV;
      double P[2][3], cg[2][3], psi[6];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int k = 0; k < 3; ++k) P[h][k] = M[k][0] * g[h][0] + M[k][1] * g[h][1] + M[k][2] * g[h][2];
        symMul(Ci, g[h], cg[h]);
      }
#pragma unroll
      for (int v = 0; v < 6; ++v) {
        const int k = vI(v), l = vJ(v);
        psi[v]      = w2d * (M[k][0] * F[l][0] + M[k][1] * F[l][1] + M[k][2] * F[l][2]);
      }
      // W terms: A = p_{alpha,k} (or l), B = w p_{beta,l} (or k)
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l) {
          K01[3 * k + l] = __builtin_amdgcn_mfma_f64_4x4x4f64(P[0][k], w1d * P[1][l], K01[3 * k + l], 0, 0, 0);
          K01[3 * k + l] = __builtin_amdgcn_mfma_f64_4x4x4f64(P[0][l], w2d * P[1][k], K01[3 * k + l], 0, 0, 0);
          if (k <= l) {
            const int v = voigt(k, l);
            K00[v]      = __builtin_amdgcn_mfma_f64_4x4x4f64(P[0][k], w1d * P[0][l], K00[v], 0, 0, 0);
            K00[v]      = __builtin_amdgcn_mfma_f64_4x4x4f64(P[0][l], w2d * P[0][k], K00[v], 0, 0, 0);
            K11[v]      = __builtin_amdgcn_mfma_f64_4x4x4f64(P[1][k], w1d * P[1][l], K11[v], 0, 0, 0);
            K11[v]      = __builtin_amdgcn_mfma_f64_4x4x4f64(P[1][l], w2d * P[1][k], K11[v], 0, 0, 0);
          }
        }
      // Psi and g^T S g terms: A = g_{alpha,I}, B = Psi_v cg_{beta,I} + d_kl sg_{beta,I}
#pragma unroll
      for (int v = 0; v < 6; ++v) {
        const bool dg = v < 3;
#pragma unroll
        for (int I = 0; I < 3; ++I) {
          const double b0 = dg ? psi[v] * cg[0][I] + sg[0][I] : psi[v] * cg[0][I];
          const double b1 = dg ? psi[v] * cg[1][I] + sg[1][I] : psi[v] * cg[1][I];
          if (dg)
            K01[4 * v] = __builtin_amdgcn_mfma_f64_4x4x4f64(g[0][I], b1, K01[4 * v], 0, 0, 0);
          else
            G01[v - 3] = __builtin_amdgcn_mfma_f64_4x4x4f64(g[0][I], b1, G01[v - 3], 0, 0, 0);
          K00[v] = __builtin_amdgcn_mfma_f64_4x4x4f64(g[0][I], b0, K00[v], 0, 0, 0);
          K11[v] = __builtin_amdgcn_mfma_f64_4x4x4f64(g[1][I], b1, K11[v], 0, 0, 0);
        }
      }
    }
  }

  if constexpr (WANT_R) {
    if (valid && i < 3) {
      const unsigned bits = fbits ? fbits[e] : 0u;
      const double t      = i == 0 ? trac.x : (i == 1 ? trac.y : trac.z);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int a = kk + 4 * h;
        double v    = Rq[h];
        if (bits) v += faceLoad(X, conn, e, bits, a, t);  // Traction skill, after the internal forces
        Rebe[(rdest ? (int64_t)rdest[e * 8 + a] * 3 : e * 24 + 3 * a) + i] = v;
      }
    }
  }
  if constexpr (WANT_K) {
#pragma unroll
    for (int v = 3; v < 6; ++v) {
      const int k = vI(v), l = vJ(v);
      K01[3 * k + l] += G01[v - 3];
      K01[3 * l + k] += G01[v - 3];
    }
    // result lane = 16 r + 4 bk + c holds rows a = r (+4), columns b = c (+4): block (r, c+4) whole and, of
    // the blocks (a, b) = (r, c) and (r+4, c+4), the entries no other lane holds. TILED: entry values go to
    // the accumulator of their group (ebe viewed as [576][9]) in the K block's orientation, rank by rank;
    // otherwise element block j = 36 s + sb is accumulator j (rank 0, as computed).
    __shared__ __align__(16) double ebe[kTileElems][324];
    {
      const int r = kk, c = i;
      double* acc = &ebe[0][0];
      const int sbW   = upperBlock(r, c + 4);
      const int sbA[2] = {r < c ? upperBlock(r, c) : (r == c ? upperBlock(r, r) : upperBlock(c, r)),
                          r < c ? upperBlock(r + 4, c + 4)
                                : (r == c ? upperBlock(r + 4, r + 4) : upperBlock(c + 4, r + 4))};
      // group | rank << 10 | transposed << 15 of the three blocks
      unsigned PW, PA[2];
      if constexpr (TILED) {
        PW    = ep[36 * s + sbW];
        PA[0] = ep[36 * s + sbA[0]];
        PA[1] = ep[36 * s + sbA[1]];
      } else {
        PW    = 36 * s + sbW;
        PA[0] = 36 * s + sbA[0];
        PA[1] = 36 * s + sbA[1];
      }
      for (int rk = 0; rk < nrank; ++rk) {
        auto put = [&](int idx, double val) {
          if (rk == 0)
            acc[idx] = val;
          else
            atomicAdd(&acc[idx], val);  // the only add to this value in this step
        };
        if (((PW >> 10) & 0xFu) == (unsigned)rk) {
          const int base = (int)(PW & 0x3FFu) * 9;
          const bool tr  = (PW >> 15) != 0;
#pragma unroll
          for (int e9 = 0; e9 < 9; ++e9) put(base + (tr ? 3 * (e9 % 3) + e9 / 3 : e9), K01[e9]);
        }
#pragma unroll
        for (int al = 0; al < 2; ++al) {
          if (((PA[al] >> 10) & 0xFu) != (unsigned)rk) continue;
          const int base = (int)(PA[al] & 0x3FFu) * 9;
          const bool tr  = (PA[al] >> 15) != 0;
#pragma unroll
          for (int v = 0; v < 6; ++v) {
            const int k = vI(v), l = vJ(v);
            const double val = al ? K11[v] : K00[v];
            // entry (k, l) of block (a, b) for r < c and of the symmetric diagonal block for r == c (with
            // its mirror), entry (l, k) of block (b, a) for r > c
            const int kl = r > c ? 3 * l + k : 3 * k + l, lk = r > c ? 3 * k + l : 3 * l + k;
            if (r <= c || k != l) put(base + (tr ? lk : kl), val);
            if (r == c && k != l) put(base + lk, val);
          }
        }
        __syncthreads();
      }
    }
    if constexpr (!TILED) {
      if (nvalidWave > 0)
        storeBlocks<TK>(&ebe[EPW * wave][0], Kebe + (e0 + (int64_t)blockIdx.x * kTileElems + EPW * wave) * 324,
                        nvalidWave * 324, lane);
    } else if (threadIdx.x < kTileThreads / 9 * 9) {
      // group sums: 9 lanes per group (entry tt = 3k + l of the 3x3 block), kTileThreads / 9 groups per
      // pass; groups come ordered by slot, so a pass writes ascending slots
      const double* acc = &ebe[0][0];
      const int tt = threadIdx.x % 9, k = tt / 3, l = tt - 3 * (tt / 3);
      for (int g = threadIdx.x / 9; g < ng; g += kTileThreads / 9) {
        double v = acc[g * 9 + tt];
        const int32_t d = gs[g];
        if (d >= 0) {
          Kebe[(int64_t)d * 9 + tt] = (TK)v;
        } else {
          // complete K block: the FULL treatment as in the gather, then the block and its mirror, stored
          // non-temporally (K is read back only by later passes, long after it has left the caches; the
          // staged sums above stay cached for the gather)
          const int32_t p = -d - 1, pt = gt[g];
          if (maskDbc) {
            const unsigned m = gm[g];
            if ((m >> tt) & 1u) v = (k == l && ((m >> (9 + k)) & 1u)) ? 1.0 : 0.0;
          }
          __builtin_nontemporal_store((TK)v, &K[(int64_t)p * 9 + tt]);
          if (pt != p) __builtin_nontemporal_store((TK)v, &K[(int64_t)pt * 9 + 3 * l + k]);
        }
      }
    }
  }
}

// Residual-only element kernel (R_e[3a+k] = sum_gp (F S dV g_a)_k): lane (el, gp) = 8 elements x 8 Gauss
// points per wavefront. The trilinear maps x(xi), u(xi) are written in monomials, x = sum_S c_S m_S(xi) with
// m_S = prod_{d in S} xi_d over the subsets S of {0,1,2} (bit d of S = dimension d, as for the corners);
// the coefficients follow from the corner values by the Moebius transform over the corner bits,
// c_S = sum_{T subset S} (-1)^|S\T| x_T: the element's lane q loads corner q and, after one exchange step
// per dimension, holds c_q. J = dx/dxi and H = du/dxi at a point then cost three FMAs per entry, and the
// point needs F = I + H J^-1 and, per corner, F S dV g_a = Q dN_a with Q = F S dV J^-T (g_a = J^-T dN_a is
// never formed). The corner forces use the same basis backwards: N_a = sum_{S superset of a} (-1)^|S\a| m_S,
// so with G_S = Q grad m_S summed over the element's points, R_a = sum_{S superset of a} (-1)^|S\a| G_S.
// The eight Gauss-point lanes reduce-scatter their G_S (exchanges at lane distance 4, 2, 1, a fixed order),
// which leaves G_q on lane q, and three more exchange steps turn G_q into corner q's three sums.
template <int KIND>
__global__ __launch_bounds__(64) void hex8Residual(const double* __restrict__ X, const double* __restrict__ U,
                                                   const int32_t* __restrict__ conn, int64_t e0, int64_t count,
                                                   double lam, double mu, const int32_t* __restrict__ rdest,
                                                   double* __restrict__ Rebe, const uint8_t* __restrict__ fbits,
                                                   double3 trac, unsigned long long* __restrict__ err) {
  __shared__ double CF[8][8][6];  // monomial coefficients of x (0..2) and u (3..5) per element
  const int lane     = threadIdx.x;
  const int el       = lane >> 3, q = lane & 7;
  const int64_t base = (int64_t)blockIdx.x * 8;
  const int nvalid   = (int)(count - base < 8 ? count - base : 8);
  const bool valid   = el < nvalid;
  const int64_t e    = e0 + base + (valid ? el : nvalid - 1);
  {
    double c[6];
    const int64_t nd = conn[e * 8 + q];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      c[k]     = X[nd * 3 + k];
      c[3 + k] = U[nd * 3 + k];
    }
    // Moebius transform: the exchange partners (lane ^ 1, 2, 4) are lanes of the same element
    auto step = [&](auto bitc) {
      constexpr int B   = decltype(bitc)::value;
      const double sub  = (q & B) ? 1.0 : 0.0;  // c - 1 * o exactly as c - o, c - 0 * o = c
#pragma unroll
      for (int k = 0; k < 6; ++k) c[k] = fma(-sub, laneXor<B>(c[k]), c[k]);
    };
    step(std::integral_constant<int, 1>{});
    step(std::integral_constant<int, 2>{});
    step(std::integral_constant<int, 4>{});
#pragma unroll
    for (int k = 0; k < 6; ++k) CF[el][q][k] = c[k];
  }
  __syncthreads();
  double xi[3];
  gaussPoint(q, xi);
  // d/dxi0 = c1 + c3 xi1 + (c5 + c7 xi1) xi2, d/dxi1 = c2 + c3 xi0 + (c6 + c7 xi0) xi2,
  // d/dxi2 = c4 + c5 xi0 + (c6 + c7 xi0) xi1
  double J[3][3], H[3][3];
  {
    const double(*cf)[6] = CF[el];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double c70 = fma(cf[7][k], xi[0], cf[6][k]);
      double* row      = k < 3 ? J[k] : H[k - 3];
      row[0]           = fma(fma(cf[7][k], xi[1], cf[5][k]), xi[2], fma(cf[3][k], xi[1], cf[1][k]));
      row[1]           = fma(c70, xi[2], fma(cf[3][k], xi[0], cf[2][k]));
      row[2]           = fma(c70, xi[1], fma(cf[5][k], xi[0], cf[4][k]));
    }
  }
  const double detJ = det3(J);
  double Ji[3][3];
  inv3(J, detJ, Ji);
  double F[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      F[i][j] = (i == j ? 1.0 : 0.0) + H[i][0] * Ji[0][j] + H[i][1] * Ji[1][j] + H[i][2] * Ji[2][j];
  double Sv[6], Ci[6], w2;
  if (!material<KIND>(lam, mu, F, Sv, Ci, w2) && valid) atomicMin(err, (unsigned long long)e);
  const double dV = fabs(detJ) * 0.125;
#pragma unroll
  for (int v = 0; v < 6; ++v) Sv[v] *= dV;
  // Q = (F S dV) J^-T: row k of F S dV is S dV times row k of F (S symmetric)
  double Q[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double row[3] = {F[k][0], F[k][1], F[k][2]};
    double T[3];
    symMul(Sv, row, T);
#pragma unroll
    for (int j = 0; j < 3; ++j) Q[k][j] = T[0] * Ji[j][0] + T[1] * Ji[j][1] + T[2] * Ji[j][2];
  }
  // G_S = Q grad m_S at this point (G_0 = 0)
  const double x01 = xi[0] * xi[1], x02 = xi[0] * xi[2], x12 = xi[1] * xi[2];
  double G[8][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    G[0][k] = 0.0;
    G[1][k] = Q[k][0];
    G[2][k] = Q[k][1];
    G[4][k] = Q[k][2];
    G[3][k] = Q[k][0] * xi[1] + Q[k][1] * xi[0];
    G[5][k] = Q[k][0] * xi[2] + Q[k][2] * xi[0];
    G[6][k] = Q[k][1] * xi[2] + Q[k][2] * xi[1];
    G[7][k] = Q[k][0] * x12 + Q[k][1] * x02 + Q[k][2] * x01;
  }
  // reduce-scatter over the element's eight lanes: keep the half of the monomials selected by this lane's
  // bit, add the partner's partials of that half
  const bool b2 = (q >> 2) & 1, b1 = (q >> 1) & 1, b0 = q & 1;
  double n4[4][3];  // monomials j + 4 b2
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      n4[j][k] = (b2 ? G[j + 4][k] : G[j][k]) + laneXor<4>(b2 ? G[j][k] : G[j + 4][k]);
  double n2[2][3];  // monomials j + 2 b1 + 4 b2
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      n2[j][k] = (b1 ? n4[j + 2][k] : n4[j][k]) + laneXor<2>(b1 ? n4[j][k] : n4[j + 2][k]);
  double r[3];  // monomial q, then corner q
#pragma unroll
  for (int k = 0; k < 3; ++k) r[k] = (b0 ? n2[1][k] : n2[0][k]) + laneXor<1>(b0 ? n2[0][k] : n2[1][k]);
  auto back = [&](auto bitc) {
    constexpr int B  = decltype(bitc)::value;
    const double sub = (q & B) ? 0.0 : 1.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) r[k] = fma(-sub, laneXor<B>(r[k]), r[k]);
  };
  back(std::integral_constant<int, 1>{});
  back(std::integral_constant<int, 2>{});
  back(std::integral_constant<int, 4>{});
  if (valid) {
    const unsigned bits = fbits ? fbits[e] : 0u;
    const double tk[3]  = {trac.x, trac.y, trac.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double acc = r[k];
      if (bits) acc += faceLoad(X, conn, e, bits, q, tk[k]);  // Traction skill, after the internal forces
      Rebe[(rdest ? (int64_t)rdest[e * 8 + q] * 3 : e * 24 + 3 * q) + k] = acc;
    }
  }
}

// K gather: thread per (upper block, entry) over the gathered upper blocks [q0, q1) of one staging chunk (in
// gather order: by storage slot, so consecutive upper blocks are mostly neighbouring rows' blocks j, whose
// slots, and whose mirrors' slots, are adjacent). The staged blocks of upper block q occupy the contiguous
// slots [segptr[q], segptr[q+1]) (ordered by element or tile), so consecutive threads read consecutive values. The entry's sum is treated with FULL (rows/columns
// of constrained DOFs zero, unit diagonal; simpleassemblers.inl:152-169) and written to the block and,
// transposed, to its mirror block below the diagonal (K and the Dirichlet treatment are symmetric).
// NT: the staged values are read once and K is written once, both non-temporally. For the tiled plain
// element at 64^3 this took the gather from 149 to 139 us and the element and residual passes that
// follow it by 12 and 7 us (no dirty K lines left behind in the caches); the EAS element's per-element
// staging gathered that way took 632 instead of 519 us at 96x96x48, so that path keeps cached accesses.
// TS: staging type, summed in TS; T: matrix type (the EAS element stages FP64 also for an FP32 matrix).
template <typename TS, typename T, bool NT>
__global__ __launch_bounds__(256) void gatherK(const int32_t* __restrict__ segptr, const int32_t* __restrict__ upBlk,
                                               const int32_t* __restrict__ upT, const uint16_t* __restrict__ upMask,
                                               int maskDbc, const TS* __restrict__ Kst, int64_t q0, int64_t q1,
                                               T* __restrict__ K) {
  const uint32_t loc = blockIdx.x * 256u + threadIdx.x;  // (q - q0) * 9 + t
  const uint32_t ql   = loc / 9u;
  if ((int64_t)ql >= q1 - q0) return;
  const int64_t q  = q0 + ql;
  const int t      = (int)(loc - ql * 9u);
  const int k      = t / 3, l = t - 3 * (t / 3);
  int32_t c        = segptr[q];
  const int32_t c1 = segptr[q + 1];
  TS acc           = TS(0);
  // segments hold one block per element (or tile) sharing the two nodes (8 at most on hexahedral meshes
  // whose nodes have 8 elements); the loads of a group of 8 are issued together, then summed in order
  for (; c < c1; c += 8) {
    TS v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const TS* src = &Kst[(int64_t)(c + j) * 9 + t];
      v[j]          = c + j < c1 ? (NT ? __builtin_nontemporal_load(src) : *src) : TS(0);
    }
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (c + j < c1) acc += v[j];
  }
  const int32_t p = upBlk[q];
  if (maskDbc) {
    const unsigned m = upMask[q];
    if ((m >> t) & 1u) acc = (k == l && ((m >> (9 + k)) & 1u)) ? TS(1) : TS(0);
  }
  const int32_t pt = upT[q];
  const T val      = (T)acc;
  if constexpr (NT) {
    __builtin_nontemporal_store(val, &K[(int64_t)p * 9 + t]);
    if (pt != p) __builtin_nontemporal_store(val, &K[(int64_t)pt * 9 + 3 * l + k]);
  } else {
    K[(int64_t)p * 9 + t] = val;
    if (pt != p) K[(int64_t)pt * 9 + 3 * l + k] = val;
  }
}

// R gather: thread per DOF over node rows [r0, r1); the node's staged corner contributions occupy the
// contiguous slots [ncptr[n], ncptr[n+1]) (ordered by element) and are summed in that order.
__global__ __launch_bounds__(256) void gatherR(const int32_t* __restrict__ ncptr, const uint8_t* __restrict__ fixed,
                                               int maskDbc, const double* __restrict__ Rst, int64_t r0, int64_t r1,
                                               double* __restrict__ R) {
  const int64_t d = r0 * 3 + (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (d >= r1 * 3) return;
  if (maskDbc && fixed[d]) {
    R[d] = 0.0;
    return;
  }
  const int64_t node = d / 3;
  const int k        = (int)(d - node * 3);
  const int32_t c1   = ncptr[node + 1];
  double acc         = 0.0;
  for (int32_t c = ncptr[node]; c < c1; c += 8) {
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = c + j < c1 ? Rst[(int64_t)(c + j) * 3 + k] : 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (c + j < c1) acc += v[j];
  }
  R[d] = acc;
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

// tiled: tangent staging over tiles [e0, e0 + count); otherwise elements [e0, e0 + count) in element order
template <typename T, int KIND, bool WK, bool WR>
void launchElem(ik_handle* h, int64_t e0, int64_t count, bool tiled, const int32_t* rdest, T* Kebe, double* Rebe,
                hipStream_t s, T* K, int mask) {
  if constexpr (!WK) {
    const unsigned grid = (unsigned)((count + 7) / 8);
    hipLaunchKernelGGL((hex8Residual<KIND>), dim3(grid), dim3(64), 0, s, h->X.p, h->u.p, h->conn.p, e0, count,
                       h->lam, h->mu, rdest, Rebe, h->faceBits.p, loadedTraction(h), h->errFlag.p);
  } else if constexpr (!WR) {
    // tangent staging over tiles (the residual instance below serves single elements in element order)
    (void)tiled;
    hipLaunchKernelGGL((hex8Element<KIND, T, WK, WR, true>), dim3((unsigned)count), dim3(kTileThreads), 0, s, h->X.p,
                       h->u.p, h->conn.p, h->tileElem.p, e0, count, h->lam, h->mu, h->tgptr.p, h->gdst.p, h->gpt.p,
                       h->gmsk.p, h->epos.p, h->tmaxr.p, K, mask, rdest, Kebe, Rebe, h->faceBits.p, loadedTraction(h),
                       h->errFlag.p);
  } else {
    const unsigned grid = (unsigned)((count + kTileElems - 1) / kTileElems);
    hipLaunchKernelGGL((hex8Element<KIND, T, WK, WR, false>), dim3(grid), dim3(kTileThreads), 0, s, h->X.p, h->u.p,
                       h->conn.p, nullptr, e0, count, h->lam, h->mu, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr, (T*)nullptr, 0, rdest, Kebe, Rebe, h->faceBits.p, loadedTraction(h), h->errFlag.p);
  }
}

// Potential energy: the hyperelastic energy sum psi |det J| / 8 over the 2x2x2 Gauss points
// (NonLinearElastic::energyFunction, nonlinearelastic.hh:403-408) plus the traction skill's
// -t.u(face centre) |J_Gamma| on every loaded face (traction.hh:71-104, 1-point face rule). One thread per
// element (grid-stride), block sums in a fixed order.
template <int KIND>
__global__ __launch_bounds__(256) void energyKernel(const double* __restrict__ X, const double* __restrict__ U,
                                                   const int32_t* __restrict__ conn, int64_t ne, double lam,
                                                   double mu, const uint8_t* __restrict__ fbits, double3 trac,
                                                   double* __restrict__ partial,
                                                   unsigned long long* __restrict__ err) {
  const double gq = 0.28867513459481288225;
  double acc      = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < ne; e += (int64_t)gridDim.x * 256) {
    double xe[8][3], ue[8][3];
    for (int a = 0; a < 8; ++a) {
      const int64_t nd = conn[e * 8 + a];
      for (int k = 0; k < 3; ++k) {
        xe[a][k] = X[nd * 3 + k];
        ue[a][k] = U[nd * 3 + k];
      }
    }
    for (int g = 0; g < 8; ++g) {
      double xi[3];
      for (int k = 0; k < 3; ++k) xi[k] = ((g >> k) & 1) ? 0.5 + gq : 0.5 - gq;
      double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int a = 0; a < 8; ++a) {
        double fb[3], dN[3];
        for (int k = 0; k < 3; ++k) fb[k] = ((a >> k) & 1) ? xi[k] : 1.0 - xi[k];
        for (int k = 0; k < 3; ++k) dN[k] = (((a >> k) & 1) ? 1.0 : -1.0) * fb[(k + 1) % 3] * fb[(k + 2) % 3];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            J[i][j] += xe[a][i] * dN[j];
            H[i][j] += ue[a][i] * dN[j];  // du/dxi
          }
      }
      const double dJ = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                        J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                        J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
      double Ji[3][3];
      Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / dJ;
      Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / dJ;
      Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / dJ;
      Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / dJ;
      Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / dJ;
      Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / dJ;
      Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / dJ;
      Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / dJ;
      Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / dJ;
      double F[3][3];  // I + (du/dxi) J^-1
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          F[i][j] = (i == j ? 1.0 : 0.0) + H[i][0] * Ji[0][j] + H[i][1] * Ji[1][j] + H[i][2] * Ji[2][j];
      double C[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
      double psi;
      if constexpr (KIND == IK_SVK) {
        const double tr = 0.5 * (C[0][0] + C[1][1] + C[2][2] - 3.0);
        double sq       = 0.0;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            const double Eij = 0.5 * (C[i][j] - (i == j ? 1.0 : 0.0));
            sq += Eij * Eij;
          }
        psi = 0.5 * lam * tr * tr + mu * sq;
      } else {
        const double detC = C[0][0] * (C[1][1] * C[2][2] - C[1][2] * C[2][1]) -
                            C[0][1] * (C[1][0] * C[2][2] - C[1][2] * C[2][0]) +
                            C[0][2] * (C[1][0] * C[2][1] - C[1][1] * C[2][0]);
        if (!(detC > 0.0)) {
          atomicMin(err, (unsigned long long)e);
          continue;
        }
        const double lnJ = 0.5 * log(detC);
        psi = 0.5 * mu * (C[0][0] + C[1][1] + C[2][2] - 3.0 - 2.0 * lnJ) + 0.5 * lam * lnJ * lnJ;
      }
      acc += psi * fabs(dJ) * 0.125;
    }
    const unsigned bits = fbits ? fbits[e] : 0u;
    for (int f = 0; f < 6 && bits; ++f) {
      if (!((bits >> f) & 1)) continue;
      const int ax = f >> 1, sd = f & 1;
      double xi[3];
      for (int k = 0; k < 3; ++k) xi[k] = k == ax ? (double)sd : 0.5;
      double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, uc[3] = {0, 0, 0};
      for (int a = 0; a < 8; ++a) {
        double fb[3], dN[3];
        for (int k = 0; k < 3; ++k) fb[k] = ((a >> k) & 1) ? xi[k] : 1.0 - xi[k];
        for (int k = 0; k < 3; ++k) dN[k] = (((a >> k) & 1) ? 1.0 : -1.0) * fb[(k + 1) % 3] * fb[(k + 2) % 3];
        const double Na = fb[0] * fb[1] * fb[2];
        for (int i = 0; i < 3; ++i) {
          uc[i] += Na * ue[a][i];
          for (int j = 0; j < 3; ++j) J[i][j] += xe[a][i] * dN[j];
        }
      }
      double t0[3], t1[3];  // the two tangent columns of J (axes other than ax, in increasing order)
      for (int i = 0; i < 3; ++i) {
        t0[i] = ax == 0 ? J[i][1] : J[i][0];
        t1[i] = ax == 2 ? J[i][1] : J[i][2];
      }
      const double c0 = t0[1] * t1[2] - t0[2] * t1[1], c1 = t0[2] * t1[0] - t0[0] * t1[2],
                   c2 = t0[0] * t1[1] - t0[1] * t1[0];
      acc -= (trac.x * uc[0] + trac.y * uc[1] + trac.z * uc[2]) * sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    }
  }
  __shared__ double red[4];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Stress at the local point xi of every element (NonLinearElastic::calculateAtImpl,
// nonlinearelastic.hh:237-270): kind 0 PK2 S, 1 Kirchhoff F S F^T, 2 Cauchy F S F^T / det F, Voigt
// [11,22,33,23,13,12] without shear factors. One thread per element.
template <int KIND>
__global__ __launch_bounds__(256) void stressKernel(const double* __restrict__ X, const double* __restrict__ U,
                                                   const int32_t* __restrict__ conn, int64_t ne, double lam,
                                                   double mu, double3 xi3, int kind, double* __restrict__ out,
                                                   unsigned long long* __restrict__ err) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= ne) return;
  const double xi[3] = {xi3.x, xi3.y, xi3.z};
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int a = 0; a < 8; ++a) {
    const int64_t nd = conn[e * 8 + a];
    double fb[3], dN[3];
    for (int k = 0; k < 3; ++k) fb[k] = ((a >> k) & 1) ? xi[k] : 1.0 - xi[k];
    for (int k = 0; k < 3; ++k) dN[k] = (((a >> k) & 1) ? 1.0 : -1.0) * fb[(k + 1) % 3] * fb[(k + 2) % 3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        J[i][j] += X[nd * 3 + i] * dN[j];
        H[i][j] += U[nd * 3 + i] * dN[j];
      }
  }
  const double dJ = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                    J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                    J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  double Ji[3][3];
  Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / dJ;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / dJ;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / dJ;
  Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / dJ;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / dJ;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / dJ;
  Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / dJ;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / dJ;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / dJ;
  double F[3][3], C[3][3], S[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      F[i][j] = (i == j ? 1.0 : 0.0) + H[i][0] * Ji[0][j] + H[i][1] * Ji[1][j] + H[i][2] * Ji[2][j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
  if constexpr (KIND == IK_SVK) {
    const double tr = 0.5 * (C[0][0] + C[1][1] + C[2][2] - 3.0);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) S[i][j] = (i == j ? lam * tr : 0.0) + mu * (C[i][j] - (i == j ? 1.0 : 0.0));
  } else {
    const double detC = C[0][0] * (C[1][1] * C[2][2] - C[1][2] * C[2][1]) -
                        C[0][1] * (C[1][0] * C[2][2] - C[1][2] * C[2][0]) +
                        C[0][2] * (C[1][0] * C[2][1] - C[1][1] * C[2][0]);
    if (!(detC > 0.0)) {
      atomicMin(err, (unsigned long long)e);
      return;
    }
    double Ci[3][3];
    Ci[0][0] = (C[1][1] * C[2][2] - C[1][2] * C[2][1]) / detC;
    Ci[0][1] = (C[0][2] * C[2][1] - C[0][1] * C[2][2]) / detC;
    Ci[0][2] = (C[0][1] * C[1][2] - C[0][2] * C[1][1]) / detC;
    Ci[1][1] = (C[0][0] * C[2][2] - C[0][2] * C[2][0]) / detC;
    Ci[1][2] = (C[0][2] * C[1][0] - C[0][0] * C[1][2]) / detC;
    Ci[2][2] = (C[0][0] * C[1][1] - C[0][1] * C[1][0]) / detC;
    Ci[1][0] = Ci[0][1];
    Ci[2][0] = Ci[0][2];
    Ci[2][1] = Ci[1][2];
    const double lnJ = 0.5 * log(detC);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) S[i][j] = mu * ((i == j ? 1.0 : 0.0) - Ci[i][j]) + lam * lnJ * Ci[i][j];
  }
  double T[3][3];
  if (kind == 0) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) T[i][j] = S[i][j];
  } else {
    const double dF = F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) -
                      F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
                      F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double t = 0.0;
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) t += F[i][k] * S[k][l] * F[j][l];
        T[i][j] = kind == 2 ? t / dF : t;
      }
  }
  double* o = out + e * 6;
  o[0] = T[0][0];
  o[1] = T[1][1];
  o[2] = T[2][2];
  o[3] = T[1][2];
  o[4] = T[0][2];
  o[5] = T[0][1];
}

template <typename T, bool WK, bool WR>
void launchElemMat(ik_handle* h, int64_t e0, int64_t count, bool tiled, const int32_t* rdest, T* Kebe,
                   double* Rebe, hipStream_t s = nullptr, T* K = nullptr, int mask = 0) {
  if (!s) s = h->stream;
  if (h->material == IK_SVK)
    launchElem<T, IK_SVK, WK, WR>(h, e0, count, tiled, rdest, Kebe, Rebe, s, K, mask);
  else
    launchElem<T, IK_NEOHOOKE, WK, WR>(h, e0, count, tiled, rdest, Kebe, Rebe, s, K, mask);
}

}  // namespace

void launchZero(double* p, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(fillKernel<double>, dim3(gridFor(n, 256)), dim3(256), 0, s, p, n, 0.0);
}
void launchZeroF(float* p, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(fillKernel<float>, dim3(gridFor(n, 256)), dim3(256), 0, s, p, n, 0.0f);
}

// K storage of the sliced-ELL layout: the padding slots must hold zeros (the SpMV reads them); every
// assembly rewrites all real blocks, so zeroing once per allocation suffices
static void ensureK(ik::DevBuf<double>& K, int64_t n, hipStream_t s) {
  if (K.p && (int64_t)K.n >= n) return;
  K.ensure((size_t)n);
  launchZero(K.p, n, s);
}
static void ensureK(ik::DevBuf<float>& K, int64_t n, hipStream_t s) {
  if (K.p && (int64_t)K.n >= n) return;
  K.ensure((size_t)n);
  launchZeroF(K.p, n, s);
}

// the flag words copied to h->herr by a completed copy: clears the device flag and throws what they report
static void raiseFlagged(ik_handle* h) {
  h->errPending                = false;
  const unsigned long long det = h->herr[0], sing = h->herr[1];
  if (det == ~0ull && sing == ~0ull) return;
  const unsigned long long reset[2] = {~0ull, ~0ull};
  IK_HIP(hipMemcpy(h->errFlag.p, reset, sizeof(reset), hipMemcpyHostToDevice));
  if (det != ~0ull) {
    h->bad_elem = (int64_t)det;
    throw Error{IK_ENEGDET, "Determinant of right Cauchy Green tensor C must be greater than zero (element " +
                                std::to_string(det) + ")"};
  }
  h->bad_elem = (int64_t)sing;
  throw Error{IK_ESINGULAR, "the enhanced-strain matrix D of element " + std::to_string(sing) +
                                " is singular (a pivot of its elimination is zero or not finite)"};
}

// Reads the error flag after the kernels enqueued so far. mayDefer (assemblies): inside a Newton loop the copy is
// only enqueued, and the loop's next synchronisation, which a host decision needs anyway, inspects it
// (finishErrCheck): each such read otherwise leaves the device idle for a host round trip.
static void checkErrFlag(ik_handle* h, bool mayDefer = false) {
  if (!h->herr) IK_HIP(hipHostMalloc((void**)&h->herr, 2 * sizeof(unsigned long long), hipHostMallocDefault));
  IK_HIP(hipMemcpyAsync(h->herr, h->errFlag.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, h->stream));
  if (mayDefer && h->deferErr) {
    h->errPending = true;
    return;
  }
  IK_HIP(hipStreamSynchronize(h->stream));
  raiseFlagged(h);
}

void finishErrCheck(ik_handle* h, bool sync) {
  if (!h->errPending) return;
  if (sync) IK_HIP(hipStreamSynchronize(h->stream));
  raiseFlagged(h);
}

void stressAt(ik_handle* h, int kind, const double xi[3], double* out) {
  require(h->neas == 0, IK_ENOTIMPL, "stresses of EAS elements are not on the device");
  require(kind >= 0 && kind <= 2, IK_EINVAL, "stress kind: IK_PK2, IK_KIRCHHOFF or IK_CAUCHY");
  const double3 x = make_double3(xi[0], xi[1], xi[2]);
  const unsigned g = (unsigned)((h->ne + 255) / 256);
  if (h->material == IK_SVK)
    hipLaunchKernelGGL(stressKernel<IK_SVK>, dim3(g), dim3(256), 0, h->stream, h->X.p, h->u.p, h->conn.p, h->ne,
                       h->lam, h->mu, x, kind, out, h->errFlag.p);
  else
    hipLaunchKernelGGL(stressKernel<IK_NEOHOOKE>, dim3(g), dim3(256), 0, h->stream, h->X.p, h->u.p, h->conn.p,
                       h->ne, h->lam, h->mu, x, kind, out, h->errFlag.p);
  checkErrFlag(h);
}

double energy(ik_handle* h) {
  require(h->neas == 0, IK_ENOTIMPL,
          "EAS element do not support any scalar calculations, i.e. they are not derivable from a potential");
  const int g     = (int)std::min<int64_t>((h->ne + 255) / 256, kReduceBlocks);
  double* partial = partialsFor(h, g);
  if (h->material == IK_SVK)
    hipLaunchKernelGGL(energyKernel<IK_SVK>, dim3(g), dim3(256), 0, h->stream, h->X.p, h->u.p, h->conn.p, h->ne,
                       h->lam, h->mu, h->faceBits.p, loadedTraction(h), partial, h->errFlag.p);
  else
    hipLaunchKernelGGL(energyKernel<IK_NEOHOOKE>, dim3(g), dim3(256), 0, h->stream, h->X.p, h->u.p, h->conn.p,
                       h->ne, h->lam, h->mu, h->faceBits.p, loadedTraction(h), partial, h->errFlag.p);
  checkErrFlag(h);
  return foldPartials(h, g);
}

// staging units [u0, u1): elements (EAS, K_e and R_e) or tiles (plain element, K_e; blocks complete in a
// tile go straight into K)
static void launchChunk(ik_handle* h, int64_t u0, int64_t u1, void* Kebe, void* K, bool fp32, int mask) {
  if (h->neas)
    easStage(h, u0, u1 - u0, true, h->bdest.p, h->rdest.p);
  else if (fp32)
    launchElemMat<float, true, false>(h, u0, u1 - u0, true, nullptr, (float*)Kebe, nullptr, h->stream, (float*)K,
                                      mask);
  else
    launchElemMat<double, true, false>(h, u0, u1 - u0, true, nullptr, (double*)Kebe, nullptr, h->stream,
                                       (double*)K, mask);
}

// the gather of staging chunk c (the upper blocks of the rows it completes)
template <typename T>
static void launchGatherK(ik_handle* h, size_t c, int mask, const T* Kebe, T* K, hipStream_t s) {
  const int64_t q0 = h->chunkQ[c], q1 = h->chunkQ[c + 1];
  if (q1 <= q0) return;  // every block of these rows was written by the element kernel
  const dim3 grid((unsigned)(((q1 - q0) * 9 + 255) / 256));
  if (h->neas)  // the EAS element stages FP64 blocks (h->ebeK64) whatever the matrix precision
    hipLaunchKernelGGL((gatherK<double, T, false>), grid, dim3(256), 0, s, h->segptr.p, h->upBlk.p, h->upT.p,
                       h->upMask.p, mask, h->ebeK64.p, q0, q1, K);
  else
    hipLaunchKernelGGL((gatherK<T, T, true>), grid, dim3(256), 0, s, h->segptr.p, h->upBlk.p, h->upT.p, h->upMask.p,
                       mask, Kebe, q0, q1, K);
}

// residual-only element pass over the mesh on stream s (the EAS pass always runs on the main stream)
static void launchResidual(ik_handle* h, hipStream_t s) {
  if (h->neas)
    easStage(h, 0, h->ne, false, nullptr, h->rdest.p);
  else
    launchElemMat<double, false, true>(h, 0, h->ne, false, h->rdest.p, (double*)nullptr, h->ebeR.p, s);
}
static void launchGatherR(ik_handle* h, int mask, int64_t r0, int64_t r1, hipStream_t s) {
  hipLaunchKernelGGL(gatherR, dim3((unsigned)(((r1 - r0) * 3 + 255) / 256)), dim3(256), 0, s, h->ncptr.p, h->fixed.p,
                     mask, h->ebeR.p, r0, r1, h->R.p);
}

// Tangent assembly. Element chunks run on the main stream; each chunk's gathers run on the second stream,
// released by an event recorded behind the chunk, so the gathers of one chunk overlap the element work of
// the next while its staged blocks are still cached. With one chunk the gather follows on the main stream.
// wantR: the element kernel also stages R_e (EAS), gathered with K. sideR: the residual pass of the plain
// element follows the K gather on the main stream (run beside the gather on the second stream it slowed
// both). In timing mode the chunks, the gathers and the residual pass run back to back on the main stream,
// each group between one pair of events, giving stand-alone durations.
template <typename T>
static void runChunks(ik_handle* h, int dbc, bool wantR, bool sideR, T* Kebe, T* K) {
  const int mask = dbc == IK_DBC_RAW ? 0 : 1;
  const bool fp32 = sizeof(T) == 4;
  const size_t nc = h->chunkElem.size() - 1;
  while (h->chunkEvents.size() < nc) {
    hipEvent_t ev;
    IK_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    h->chunkEvents.push_back(ev);
  }
  if (h->timing) {
    {
      ScopedKernelTimer t(h, KASM_KR, h->stream, (int64_t)nc);
      for (size_t c = 0; c < nc; ++c) launchChunk(h, h->chunkElem[c], h->chunkElem[c + 1], Kebe, K, fp32, mask);
    }
    {
      ScopedKernelTimer t(h, KGATHER, h->stream, (int64_t)nc);
      for (size_t c = 0; c < nc; ++c) launchGatherK<T>(h, c, mask, Kebe, K, h->stream);
    }
    if (sideR) {
      ScopedKernelTimer t(h, KASM_R, h->stream);
      launchResidual(h, h->stream);
    }
    if (wantR || sideR) launchGatherR(h, mask, 0, h->nn, h->stream);
    return;
  }
  if (nc == 1) {
    // one chunk: the K gather follows on the main stream
    launchChunk(h, h->chunkElem[0], h->chunkElem[1], Kebe, K, fp32, mask);
    launchGatherK<T>(h, 0, mask, Kebe, K, h->stream);
    if (sideR) launchResidual(h, h->stream);
    if (wantR || sideR) launchGatherR(h, mask, 0, h->nn, h->stream);
    return;
  }
  for (size_t c = 0; c < nc; ++c) {
    launchChunk(h, h->chunkElem[c], h->chunkElem[c + 1], Kebe, K, fp32, mask);
    const int64_t r0 = h->chunkRow[c], r1 = h->chunkRow[c + 1];
    if (r1 <= r0) continue;
    IK_HIP(hipEventRecord(h->chunkEvents[c], h->stream));
    IK_HIP(hipStreamWaitEvent(h->stream2, h->chunkEvents[c], 0));
    launchGatherK<T>(h, c, mask, Kebe, K, h->stream2);
    if (wantR) launchGatherR(h, mask, r0, r1, h->stream2);
  }
  if (sideR) launchResidual(h, h->stream);  // behind the last element chunk
  // join: later work on the main stream sees every gathered row
  IK_HIP(hipEventRecord(h->joinEvent, h->stream2));
  IK_HIP(hipStreamWaitEvent(h->stream, h->joinEvent, 0));
  if (sideR) launchGatherR(h, mask, 0, h->nn, h->stream);
}

void assembleKR(ik_handle* h, int dbc, int precision, bool wantK, bool wantR) {
  const int mask = dbc == IK_DBC_RAW ? 0 : 1;
  h->ebeR.ensure((size_t)h->ne * 24);  // EAS elements always produce R_e
  if (h->neas) {
    // the EAS kernel computes the condensed R_e with K_e (in FP64; an FP32 matrix is gathered from the FP64
    // staging); both gathers run when K is wanted
    wantR = true;
  }
  // EAS: one element pass yields K_e and R_e. Plain element: the register-bound tangent kernel and the
  // residual kernel stay separate (R accumulated in the tangent kernel costs it more registers than the
  // residual kernel takes time); with both wanted they share one launch sequence and one error check
  const bool easBoth = h->neas && wantK;
  const bool sideR   = !h->neas && wantK && wantR;
  if (wantK && precision == IK_FP64) {
    ensureK(h->K64, h->nsell * 9, h->stream);
    h->ebeK64.ensure((size_t)h->nstage * 9);
    runChunks<double>(h, dbc, easBoth, sideR, h->ebeK64.p, h->K64.p);
  } else if (wantK) {
    ensureK(h->K32, h->nsell * 9, h->stream);
    if (h->neas) {
      h->ebeK64.ensure((size_t)h->nstage * 9);
      runChunks<float>(h, dbc, easBoth, sideR, nullptr, h->K32.p);
    } else {
      h->ebeK32.ensure((size_t)h->nstage * 9);
      runChunks<float>(h, dbc, false, sideR, h->ebeK32.p, h->K32.p);
    }
  }
  if (wantR && !wantK) {
    // residual alone: one element launch over the mesh, then one gather (a few MB; no staging overlap needed)
    {
      ScopedKernelTimer t(h, KASM_R, h->stream);
      launchResidual(h, h->stream);
    }
    launchGatherR(h, mask, 0, h->nn, h->stream);
  }
  if (h->neas) h->easCacheValid = true;
  if (wantR && h->npl) {
    hipLaunchKernelGGL(pointLoadKernel, dim3(gridFor(h->npl, 128)), dim3(128), 0, h->stream, h->plDof.p,
                       h->plCoef.p, h->npl, h->lambda, h->R.p);
  }
  checkErrFlag(h, true);
}

void elementHost(ik_handle* h, int64_t e, double* Ke, double* Re) {
  require(e >= 0 && e < h->ne, IK_EINVAL, "element index out of range");
  h->ebeK64.ensure((size_t)h->ne * 324);
  h->ebeR.ensure((size_t)h->ne * 24);
  if (h->neas)
    easStage(h, e, 1, true, nullptr, nullptr);
  else
    launchElemMat<double, true, true>(h, e, 1, false, nullptr, h->ebeK64.p, h->ebeR.p);
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
