// mlp.hip — the learned load-step predictor: a per-node MLP 14 -> 128 -> 128 -> 128 -> 3 with SiLU
// activations, trained and evaluated on the GPU. No reference counterpart: it replaces the constant
// (loadcontrol.inl:55-57) / secant (pathfollowingfunctions.hh:188-195) displacement predictor, and
// NewtonRaphson still has to reach the same ||R|| tolerance from the predicted state.
//
// Training products run through one tiled GEMM on v_mfma_f32_16x16x16bf16_1k: operands are rounded to BF16
// (round to nearest even) while staged into LDS, accumulation is FP32, parameters keep an FP32 master copy
// updated by Adam. Strided operand access covers every product of training: Z = A W^T + b (forward),
// dW = dZ^T A (weight gradient, split over the batch), dA = dZ W (back-propagation). The prediction inside the
// load loop runs in one fused kernel (features, four layers, update of u) on the same instruction and rounding.
//
// Features of node v for the step n -> n+1 (load increment dlam, velocities v_n = du_n / dlam):
//   [ X_v / L (3), t / 40 (3), lam_{n+1}, 40 dlam, v_n / 0.1 (3), v_{n-1} / 0.1 (3) ]
// Output y (3): v_{n+1} = v_n + 0.1 y, u_pred = u_n + dlam v_{n+1} (clamped DOFs keep u_n).
#include <cmath>
#include <mutex>

#include "ik_internal.hh"

namespace ik {

constexpr int kIn = 14, kHid = 128, kOut = 3;
constexpr int64_t kParams = kHid * kIn + kHid + 2 * (kHid * kHid + kHid) + kOut * kHid + kOut;
// parameter offsets
constexpr int64_t oW1 = 0, oB1 = oW1 + kHid * kIn, oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid, oW3 = oB2 + kHid,
                  oB3 = oW3 + kHid * kHid, oW4 = oB3 + kHid, oB4 = oW4 + kOut * kHid;
constexpr double kVelScale = 0.1, kTracScale = 1.0 / 40.0, kDlamScale = 40.0;

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// FP32 -> BF16 bits, round to nearest even (NaN kept quiet): v_cvt_pk_bf16_f32
__device__ inline uint16_t toBf16(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

// logistic function and SiLU; the reciprocal is v_rcp_f32 (1 ulp), 1 + e^-z >= 1 so it never sees a denormal
__device__ inline float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.f + __expf(-z)); }
__device__ inline float silu(float z) { return z * sigmoid(z); }

// EPI_SILU_ONLY: C = silu(A B + bias), the pre-activation is not kept (inference)
// EPI_PARTIAL: slice z of a K split stores its partial product at C + z * M * ldc (summed later in z order)
enum { EPI_STORE = 0, EPI_PARTIAL = 1, EPI_BIAS_SILU = 2, EPI_BIAS = 3, EPI_SILU_ONLY = 4 };

// C[i][j] = sum_k A(i,k) B(k,j) with A(i,k) = A[i*sai + k*sak], B(k,j) = B[k*sbk + j*sbj].
// 256 threads, 64x64 block tile, 4 waves of 32x32 (2x2 MFMA tiles), K in steps of 16.
// gridDim.z splits K into chunks of kChunk (EPI_PARTIAL: one partial product per chunk).
template <int EPI>
__global__ __launch_bounds__(256) void gemmBf16(int M, int N, int K, const float* __restrict__ A, int64_t sai,
                                                int64_t sak, const float* __restrict__ B, int64_t sbk, int64_t sbj,
                                                float* __restrict__ C, int64_t ldc, const float* __restrict__ bias,
                                                float* __restrict__ act, int kChunk) {
  __shared__ uint16_t As[64][20];  // row i, k (padded)
  __shared__ uint16_t Bs[16][68];  // k, col j (padded)
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
  const int kb = blockIdx.z * kChunk, ke = min(K, kb + kChunk);
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += 16) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid * 4 + q;
      const int r = e >> 4, kk = e & 15;
      const int gi = i0 + r, gk = k0 + kk;
      As[r][kk] = toBf16((gi < M && gk < ke) ? A[(int64_t)gi * sai + (int64_t)gk * sak] : 0.f);
      const int kb2 = e >> 6, cc = e & 63;
      const int gk2 = k0 + kb2, gj = j0 + cc;
      Bs[kb2][cc] = toBf16((gj < N && gk2 < ke) ? B[(int64_t)gk2 * sbk + (int64_t)gj * sbj] : 0.f);
    }
    __syncthreads();
#pragma unroll
    for (int mi = 0; mi < 2; ++mi) {
      bf16x4 av;
      const int ar = wm * 32 + mi * 16 + (lane & 15);
#pragma unroll
      for (int q = 0; q < 4; ++q) av[q] = (short)As[ar][4 * (lane >> 4) + q];
#pragma unroll
      for (int nj = 0; nj < 2; ++nj) {
        bf16x4 bv;
        const int bc = wn * 32 + nj * 16 + (lane & 15);
#pragma unroll
        for (int q = 0; q < 4; ++q) bv[q] = (short)Bs[4 * (lane >> 4) + q][bc];
        acc[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[mi][nj], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int nj = 0; nj < 2; ++nj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wm * 32 + mi * 16 + 4 * (lane >> 4) + r;
        const int j = j0 + wn * 32 + nj * 16 + (lane & 15);
        if (i >= M || j >= N) continue;
        float v = acc[mi][nj][r];
        if constexpr (EPI == EPI_PARTIAL) {
          C[(int64_t)blockIdx.z * M * ldc + (int64_t)i * ldc + j] = v;
        } else {
          if constexpr (EPI == EPI_BIAS_SILU || EPI == EPI_BIAS || EPI == EPI_SILU_ONLY) v += bias[j];
          if constexpr (EPI == EPI_SILU_ONLY) {
            C[(int64_t)i * ldc + j] = silu(v);
          } else {
            C[(int64_t)i * ldc + j] = v;
          }
          if constexpr (EPI == EPI_BIAS_SILU) act[(int64_t)i * ldc + j] = silu(v);
        }
      }
}

// dZ = dA * silu'(Z)
__global__ void siluBackward(const float* __restrict__ dA, const float* __restrict__ Z, float* __restrict__ dZ,
                             int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const float z = Z[i], sg = sigmoid(z);
    dZ[i]         = dA[i] * sg * (1.f + z * (1.f - sg));
  }
}

// dY = 2 (Y - T) / (3 n); per-block partial sums of (Y - T)^2
__global__ void mseGrad(const float* __restrict__ Y, const float* __restrict__ T, int64_t n3, float scale,
                        float* __restrict__ dY, double* __restrict__ partial) {
  __shared__ double red[4];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n3; i += (int64_t)gridDim.x * 256) {
    const float d = Y[i] - T[i];
    dY[i]         = 2.f * d * scale;
    s += (double)d * d;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// db[j] += sum over rows of dZ (one block per 256-row slab, one thread per column)
// column sums of 256-row slices of dZ: part[slice * cols + j]
__global__ void biasGrad(const float* __restrict__ dZ, int64_t n, int cols, float* __restrict__ part) {
  const int j = threadIdx.x;
  if (j >= cols) return;
  const int64_t r0 = (int64_t)blockIdx.x * 256, r1 = min<int64_t>(n, r0 + 256);
  float s = 0.f;
  for (int64_t r = r0; r < r1; ++r) s += dZ[r * cols + j];
  part[(int64_t)blockIdx.x * cols + j] = s;
}

// out[i] = sum over z = 0 .. nz-1 of part[z * len + i], in that order: the gradients do not depend on scheduling
__global__ void foldSlices(const float* __restrict__ part, int nz, int64_t len, float* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < len; i += (int64_t)gridDim.x * 256) {
    float s = 0.f;
    for (int z = 0; z < nz; ++z) s += part[(int64_t)z * len + i];
    out[i] = s;
  }
}

__global__ void adamKernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float c1,
                           float c2) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i]           = mi;
    v[i]           = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  }
}

__global__ void fillF(float* __restrict__ p, int64_t n, float v) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = v;
}

// BF16 copy of the network for the fused predictor, row-major output by input: W1 [128][16] (inputs past 14
// zero), W2 and W3 [128][kWs], W4 [16][kWs] (rows past the 3 outputs zero). The rows of 128 inputs are padded to
// kWs = 136 values (68 dwords), so the 16 rows one MFMA A fragment reads start on 16 different bank quadruples
// instead of all on bank 0; the whole copy (kWbN values, 78 KB) is staged into LDS as it lies. Within each group
// of 32 inputs 32 s .. 32 s + 31 the columns are stored in the order the 16x16x32 MFMA consumes them: lane group
// g (lanes 16 g .. 16 g + 15) takes slots 8 g .. 8 g + 7 = inputs 32 s + 4 g + (0..3) and 32 s + 16 + 4 g + (0..3).
constexpr int kWs = kHid + 8;
constexpr int oWb1 = 0, oWb2 = oWb1 + kHid * 16, oWb3 = oWb2 + kHid * kWs, oWb4 = oWb3 + kHid * kWs,
              kWbN = oWb4 + 16 * kWs;
static_assert(kWbN % 8 == 0, "staged in 16-byte pieces");

__global__ void packBf16Weights(const float* __restrict__ P, uint16_t* __restrict__ Wb) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kWbN) return;
  float v = 0.f;
  if (i < oWb2) {
    const int j = i / 16, k = i % 16;
    v = k < kIn ? P[oW1 + j * kIn + k] : 0.f;
  } else {
    const int l = i < oWb3 ? 0 : i < oWb4 ? 1 : 2;
    const int o = i - (l == 0 ? oWb2 : l == 1 ? oWb3 : oWb4);
    const int j = o / kWs, c = o % kWs;
    if (c < kHid && (l < 2 || j < kOut)) {
      const int g = (c % 32) / 8, q = c % 8;
      const int k = c / 32 * 32 + (q < 4 ? 4 * g + q : 16 + 4 * g + q - 4);
      v           = P[(l == 0 ? oW2 : l == 1 ? oW3 : oW4) + j * kHid + k];
    }
  }
  Wb[i] = toBf16(v);
}

// four (eight) consecutive BF16 values of a weight row in LDS, 8- (16-) byte aligned: an MFMA A fragment
__device__ inline bf16x4 wfrag(const uint16_t* w) { return *reinterpret_cast<const bf16x4*>(w); }
__device__ inline bf16x8 wfrag8(const uint16_t* w) { return *reinterpret_cast<const bf16x8*>(w); }
// the 16x16x32 B fragment of K step s: this lane's outputs of tiles 2 s and 2 s + 1
__device__ inline bf16x8 actPair(bf16x4 lo, bf16x4 hi) {
  return __builtin_bit_cast(bf16x8, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
}

// The in-solver prediction in one launch: the per-node features (file header; velocities from the two previous
// increments), the three SiLU layers, the output layer and u <- u + dlam (v_n + 0.1 y) on free DOFs, with the
// activations kept in registers. A wavefront takes 16 nodes at a time and computes each layer transposed,
// out^T (128 x 16) = W (128 x 128) act^T, with W as the MFMA A operand (from LDS) and the activations as the B
// operand. Lane (n, kq) = (lane & 15, lane >> 4) holds node n's outputs 16 t + 4 kq + r of output tile t (C
// fragment). Layer 1 (K = 16) runs on v_mfma_f32_16x16x16_bf16; layers 2-4 on v_mfma_f32_16x16x32_bf16, whose K
// step s takes from this lane exactly its outputs of tiles 2 s and 2 s + 1 (the weight columns are stored in
// that order): no data moves between lanes from layer to layer. The 8 output tiles of a layer are independent
// accumulation chains, interleaved per K step. Same operand rounding (BF16, round to nearest even), FP32
// accumulation, bias and SiLU as the layer-by-layer GEMMs (within a 32-input step the summation order differs).
__global__ __launch_bounds__(256) void mlpPredictFused(const uint16_t* __restrict__ Wb, const float* __restrict__ P,
                                                       const double* __restrict__ X, const double* __restrict__ duN,
                                                       const double* __restrict__ duN1, int64_t nn, double invL,
                                                       double t0, double t1, double t2, double lamNext, double dlam,
                                                       double dlamPrev, double dlamPrev2, double* __restrict__ u,
                                                       const uint8_t* __restrict__ fixed) {
  __shared__ __align__(16) uint16_t sW[kWbN];  // 78 KB
  __shared__ float sB[3 * kHid + 16];
  for (int i = threadIdx.x; i < kWbN / 8; i += 256)
    reinterpret_cast<uint4*>(sW)[i] = reinterpret_cast<const uint4*>(Wb)[i];
  for (int i = threadIdx.x; i < 3 * kHid + 16; i += 256) {
    float v = 0.f;
    if (i < kHid) v = P[oB1 + i];
    else if (i < 2 * kHid) v = P[oB2 + i - kHid];
    else if (i < 3 * kHid) v = P[oB3 + i - 2 * kHid];
    else if (i - 3 * kHid < kOut) v = P[oB4 + i - 3 * kHid];
    sB[i] = v;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, n = lane & 15, kq = lane >> 4;
  const int64_t nblk = (nn + 15) / 16;
  for (int64_t blk = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); blk < nblk; blk += (int64_t)gridDim.x * 4) {
    const int64_t v = blk * 16 + n;
    const bool valid = v < nn;
    // input fragment: features 4 kq .. 4 kq + 3 of node v (as predictorFeatures)
    float f[4] = {0.f, 0.f, 0.f, 0.f};
    if (valid) {
      if (kq == 0) {
        f[0] = (float)(X[v * 3] * invL);
        f[1] = (float)(X[v * 3 + 1] * invL);
        f[2] = (float)(X[v * 3 + 2] * invL);
        f[3] = (float)(t0 * kTracScale);
      } else if (kq == 1) {
        f[0] = (float)(t1 * kTracScale);
        f[1] = (float)(t2 * kTracScale);
        f[2] = (float)lamNext;
        f[3] = (float)(dlam * kDlamScale);
      } else if (kq == 2) {
        for (int k = 0; k < 3; ++k) f[k] = dlamPrev > 0 ? (float)(duN[v * 3 + k] / dlamPrev / kVelScale) : 0.f;
        f[3] = dlamPrev2 > 0 ? (float)(duN1[v * 3] / dlamPrev2 / kVelScale) : 0.f;
      } else {
        for (int k = 0; k < 2; ++k) f[k] = dlamPrev2 > 0 ? (float)(duN1[v * 3 + 1 + k] / dlamPrev2 / kVelScale) : 0.f;
      }
    }
    bf16x4 act[8];
    act[0] = bf16x4{(short)toBf16(f[0]), (short)toBf16(f[1]), (short)toBf16(f[2]), (short)toBf16(f[3])};
    f32x4 acc[8];
    // layer 1: K = 16 (one step)
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const bf16x4 a = wfrag(sW + oWb1 + (16 * t + n) * 16 + 4 * kq);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, act[0], f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
    }
    // bias, SiLU, BF16: the next layer's B fragments
    auto activate = [&](const float* bias) {
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        short q[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) q[r] = (short)toBf16(silu(acc[t][r] + bias[16 * t + 4 * kq + r]));
        act[t] = bf16x4{q[0], q[1], q[2], q[3]};
      }
    };
    activate(sB);
    // layers 2 and 3: 4 K steps of 32 x 8 independent output tiles
#pragma unroll
    for (int l = 0; l < 2; ++l) {
      const uint16_t* W = sW + (l == 0 ? oWb2 : oWb3);
#pragma unroll
      for (int t = 0; t < 8; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const bf16x8 b = actPair(act[2 * ks], act[2 * ks + 1]);
#pragma unroll
        for (int t = 0; t < 8; ++t)
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wfrag8(W + (16 * t + n) * kWs + 32 * ks + 8 * kq), b,
                                                           acc[t], 0, 0, 0);
      }
      activate(sB + (l + 1) * kHid);
    }
    // output layer: one tile (rows 0..2 of 16 used), 4 K steps of 32
    f32x4 y = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
      y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wfrag8(sW + oWb4 + n * kWs + 32 * ks + 8 * kq),
                                                  actPair(act[2 * ks], act[2 * ks + 1]), y, 0, 0, 0);
    // lanes kq = 0 hold outputs 0..3 of node n: u <- u + dlam (v_n + 0.1 y) on free DOFs
    if (valid && kq == 0) {
#pragma unroll
      for (int k = 0; k < kOut; ++k) {
        const int64_t i = v * 3 + k;
        if (fixed[i]) continue;
        const float yk   = y[k] + sB[3 * kHid + k];
        const double vn  = dlamPrev > 0 ? duN[i] / dlamPrev : 0.0;
        u[i] += dlam * (vn + kVelScale * (double)yk);
      }
    }
  }
}

inline unsigned g1(int64_t n) { return (unsigned)std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 4096); }

}  // namespace

// ------------------------------------------------------------------------------------------------
// MLP context shared by the training entry points and the in-solver predictor
// ------------------------------------------------------------------------------------------------
struct Mlp {
  hipStream_t s = nullptr;
  int64_t maxBatch = 0;
  DevBuf<float> P, G, Mo, Vo;          // params (fp32 master), grads, Adam moments
  DevBuf<float> Z[3], Aa[3], Y, dY, dA, dZ, Xin, Tin;
  DevBuf<float> Gpart;  // per-slice partial gradients, folded in a fixed order
  DevBuf<uint16_t> Wb;  // BF16 weights of the fused predictor (packed when the parameters are set)
  DevBuf<double> partial;
  double* hred = nullptr;
  int64_t step = 0;
  std::string err;

  void ensure(int64_t n) {
    if (n <= maxBatch) return;
    maxBatch = n;
    for (int l = 0; l < 3; ++l) {
      Z[l].ensure(n * kHid);
      Aa[l].ensure(n * kHid);
    }
    Y.ensure(n * kOut);
    dY.ensure(n * kOut);
    dA.ensure(n * kHid);
    dZ.ensure(n * kHid);
    Xin.ensure(n * kIn);
    Tin.ensure(n * kOut);
  }

  // returns the number of K slices (gridDim.z)
  template <int EPI>
  int gemm(int M, int N, int K, const float* A, int64_t sai, int64_t sak, const float* B, int64_t sbk, int64_t sbj,
           float* C, int64_t ldc, const float* bias, float* act, int splitK = 1) {
    const int kChunk = splitK > 1 ? ((K + splitK - 1) / splitK + 15) / 16 * 16 : K;
    const int nz     = splitK > 1 ? (K + kChunk - 1) / kChunk : 1;
    dim3 grid((N + 63) / 64, (M + 63) / 64, nz);
    hipLaunchKernelGGL((gemmBf16<EPI>), grid, dim3(256), 0, s, M, N, K, A, sai, sak, B, sbk, sbj, C, ldc, bias, act,
                       kChunk);
    return nz;
  }

  // forward on x (n x 14) -> Y (n x 3), keeping Z/A for the backward pass
  void forward(const float* x, int64_t n) {
    const float* P0 = P.p;
    gemm<EPI_BIAS_SILU>((int)n, kHid, kIn, x, kIn, 1, P0 + oW1, 1, kIn, Z[0].p, kHid, P0 + oB1, Aa[0].p);
    gemm<EPI_BIAS_SILU>((int)n, kHid, kHid, Aa[0].p, kHid, 1, P0 + oW2, 1, kHid, Z[1].p, kHid, P0 + oB2, Aa[1].p);
    gemm<EPI_BIAS_SILU>((int)n, kHid, kHid, Aa[1].p, kHid, 1, P0 + oW3, 1, kHid, Z[2].p, kHid, P0 + oB3, Aa[2].p);
    gemm<EPI_BIAS>((int)n, kOut, kHid, Aa[2].p, kHid, 1, P0 + oW4, 1, kHid, Y.p, kOut, P0 + oB4, nullptr);
  }

  // the same forward through caller-provided buffers: hidden activations ping-pong between a and b (n x 128
  // each), the output goes to y (n x 3); no pre-activations are kept and nothing is allocated
  void forwardInto(const float* x, int64_t n, float* a, float* b, float* y) {
    const float* P0 = P.p;
    gemm<EPI_SILU_ONLY>((int)n, kHid, kIn, x, kIn, 1, P0 + oW1, 1, kIn, a, kHid, P0 + oB1, nullptr);
    gemm<EPI_SILU_ONLY>((int)n, kHid, kHid, a, kHid, 1, P0 + oW2, 1, kHid, b, kHid, P0 + oB2, nullptr);
    gemm<EPI_SILU_ONLY>((int)n, kHid, kHid, b, kHid, 1, P0 + oW3, 1, kHid, a, kHid, P0 + oB3, nullptr);
    gemm<EPI_BIAS>((int)n, kOut, kHid, a, kHid, 1, P0 + oW4, 1, kHid, y, kOut, P0 + oB4, nullptr);
  }

  // mean-squared-error loss and its gradient w.r.t. all parameters (into G)
  double lossGrad(const float* x, const float* t, int64_t n) {
    forward(x, n);
    const int nb = (int)std::min<int64_t>((n * kOut + 255) / 256, 1024);
    partial.ensure(1024);
    hipLaunchKernelGGL(mseGrad, dim3(nb), dim3(256), 0, s, Y.p, t, n * kOut, (float)(1.0 / (double)(n * kOut)), dY.p,
                       partial.p);
    const int split = (int)std::max<int64_t>(1, std::min<int64_t>(n / 256, 256));
    const unsigned nbias = (unsigned)((n + 255) / 256);
    Gpart.ensure(std::max<int64_t>((int64_t)split * kHid * kHid, (int64_t)nbias * kHid));
    // dW (M x N) = dZ^T A over the batch: one partial product per batch slice, folded in slice order
    auto weightGrad = [&](int M, int N, const float* dz, const float* ain, float* dst) {
      const int nz = gemm<EPI_PARTIAL>(M, N, (int)n, dz, 1, M, ain, N, 1, Gpart.p, N, nullptr, nullptr, split);
      hipLaunchKernelGGL(foldSlices, dim3(g1((int64_t)M * N)), dim3(256), 0, s, Gpart.p, nz, (int64_t)M * N, dst);
    };
    auto biasGradient = [&](const float* dz, int cols, float* dst) {
      hipLaunchKernelGGL(biasGrad, dim3(nbias), dim3(kHid), 0, s, dz, n, cols, Gpart.p);
      hipLaunchKernelGGL(foldSlices, dim3(g1(cols)), dim3(256), 0, s, Gpart.p, (int)nbias, (int64_t)cols, dst);
    };
    // layer 4: dW4 = dY^T A3, db4, dA3 = dY W4
    weightGrad(kOut, kHid, dY.p, Aa[2].p, G.p + oW4);
    biasGradient(dY.p, kOut, G.p + oB4);
    gemm<EPI_STORE>((int)n, kHid, kOut, dY.p, kOut, 1, P.p + oW4, kHid, 1, dA.p, kHid, nullptr, nullptr);
    for (int l = 2; l >= 0; --l) {
      hipLaunchKernelGGL(siluBackward, dim3(g1(n * kHid)), dim3(256), 0, s, dA.p, Z[l].p, dZ.p, n * kHid);
      const float* ain = l == 0 ? x : Aa[l - 1].p;
      const int kin    = l == 0 ? kIn : kHid;
      const int64_t oW = l == 0 ? oW1 : (l == 1 ? oW2 : oW3), oB = l == 0 ? oB1 : (l == 1 ? oB2 : oB3);
      weightGrad(kHid, kin, dZ.p, ain, G.p + oW);
      biasGradient(dZ.p, kHid, G.p + oB);
      if (l > 0) gemm<EPI_STORE>((int)n, kHid, kHid, dZ.p, kHid, 1, P.p + oW, kin, 1, dA.p, kHid, nullptr, nullptr);
    }
    IK_HIP(hipMemcpyAsync(hred, partial.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
    IK_HIP(hipStreamSynchronize(s));
    double sum = 0.0;
    for (int i = 0; i < nb; ++i) sum += hred[i];
    return sum / (double)(n * kOut);
  }

  void adam(float lr, float b1, float b2, float eps) {
    ++step;
    const float c1 = 1.f / (1.f - std::pow(b1, (float)step)), c2 = 1.f / (1.f - std::pow(b2, (float)step));
    hipLaunchKernelGGL(adamKernel, dim3(g1(kParams)), dim3(256), 0, s, P.p, G.p, Mo.p, Vo.p, kParams, lr, b1, b2, eps,
                       c1, c2);
  }
};

Mlp* mlpCreate(hipStream_t s) {
  auto* m = new Mlp;
  m->s    = s;
  m->P.ensure(kParams);
  m->G.ensure(kParams);
  m->Mo.ensure(kParams);
  m->Vo.ensure(kParams);
  hipLaunchKernelGGL(fillF, dim3(g1(kParams)), dim3(256), 0, s, m->Mo.p, kParams, 0.f);
  hipLaunchKernelGGL(fillF, dim3(g1(kParams)), dim3(256), 0, s, m->Vo.p, kParams, 0.f);
  IK_HIP(hipHostMalloc((void**)&m->hred, 1024 * sizeof(double), hipHostMallocDefault));
  return m;
}

void mlpDestroy(Mlp* m) {
  if (!m) return;
  if (m->s) (void)hipStreamSynchronize(m->s);
  if (m->hred) (void)hipHostFree(m->hred);
  delete m;
}

int64_t mlpNumParams() { return kParams; }

void mlpSetParams(Mlp* m, const float* p, int where, hipStream_t s) {
  IK_HIP(hipMemcpyAsync(m->P.p, p, kParams * sizeof(float),
                        where == IK_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
  m->Wb.ensure(kWbN);
  hipLaunchKernelGGL(packBf16Weights, dim3((kWbN + 255) / 256), dim3(256), 0, s, m->P.p, m->Wb.p);
  IK_HIP(hipStreamSynchronize(s));
}

void mlpGetParams(Mlp* m, float* p, int where, hipStream_t s) {
  IK_HIP(hipMemcpyAsync(p, m->P.p, kParams * sizeof(float),
                        where == IK_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  IK_HIP(hipStreamSynchronize(s));
}

// in-solver prediction: u (device) <- MLP prediction for lamNext from the two previous increments
void mlpPredict(ik_handle* h, const double* duN, const double* duN1, double lamNext, double dlam, double dlamPrev,
                double dlamPrev2) {
  require(h->mlp != nullptr && h->mlpLoaded, IK_EINVAL, "MLP predictor selected but no parameters set");
  Mlp* m = h->mlp;
  m->s   = h->stream;
  ScopedKernelTimer t(h, KMLP);
  // one fused launch; 2 workgroups of 80 KB LDS per CU, each wavefront looping over blocks of 16 nodes
  const int64_t blocks = (h->nn + 63) / 64;
  const unsigned grid  = (unsigned)std::max<int64_t>(1, std::min<int64_t>(blocks, 512));
  hipLaunchKernelGGL(mlpPredictFused, dim3(grid), dim3(256), 0, h->stream, m->Wb.p, m->P.p, h->X.p, duN, duN1, h->nn,
                     1.0 / h->mlpCharLen, h->mlpTraction[0], h->mlpTraction[1], h->mlpTraction[2], lamNext, dlam,
                     dlamPrev, dlamPrev2, h->u.p, h->fixed.p);
}

}  // namespace ik

// ------------------------------------------------------------------------------------------------
// C ABI of the training context
// ------------------------------------------------------------------------------------------------
struct ik_mlp {
  ik::Mlp* m = nullptr;
  hipStream_t s = nullptr;
  int device = 0;            // the device current at ik_mlp_create: the context's stream and memory live there
  std::string err;
};

// Scratch of the allocation-free forward (ik_mlp_forward_noalloc): two ping-pong buffers of hidden activations,
// IK_MLP_NOALLOC_MAX_ROWS x 128 FP32 each (8 MiB), one copy per device held until the process ends. Host code
// uses only the address hipGetSymbolAddress returns for the device (the array's name on the host is not a device
// address), resolved once per device at ik_mlp_create with that device current.
static __device__ __align__(256) float scratch[2 * IK_MLP_NOALLOC_MAX_ROWS * ik::kHid];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
float* g_scratchPtr[kMaxDevices] = {};

float* resolveScratch(int device) {
  ik::require(device >= 0 && device < kMaxDevices, IK_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (!g_scratchPtr[device]) {
    void* ptr = nullptr;
    IK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratchPtr[device] = static_cast<float*>(ptr);
  }
  return g_scratchPtr[device];
}
}  // namespace

namespace {
template <class F>
int mguard(ik_mlp* m, F&& f) {
  try {
    f();
    if (m) m->err.clear();
    return IK_OK;
  } catch (const ik::Error& e) {
    if (m) m->err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    if (m) m->err = e.what();
    return IK_EINVAL;
  }
}
// He-uniform weights U(-sqrt(6/fan_in), +) from SplitMix64, zero biases, last layer scaled by 0.1
void initParams(std::vector<float>& p, uint64_t seed) {
  using namespace ik;
  p.assign(kParams, 0.f);
  uint64_t st = seed;
  auto uni = [&]() {
    uint64_t z = (st += 0x9E3779B97F4A7C15ULL);
    z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z          = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int64_t off, int64_t count, int fanIn, double scale) {
    const double b = std::sqrt(6.0 / fanIn) * scale;
    for (int64_t i = 0; i < count; ++i) p[off + i] = (float)((2.0 * uni() - 1.0) * b);
  };
  fill(oW1, kHid * kIn, kIn, 1.0);
  fill(oW2, kHid * kHid, kHid, 1.0);
  fill(oW3, kHid * kHid, kHid, 1.0);
  fill(oW4, kOut * kHid, kHid, 0.1);
}
}  // namespace

extern "C" {

int ik_mlp_create(ik_mlp** out, uint64_t seed) {
  if (!out) return IK_EINVAL;
  auto* m = new ik_mlp;
  *out    = m;
  return mguard(m, [&] {
    IK_HIP(hipGetDevice(&m->device));
    (void)resolveScratch(m->device);  // the device's scratch address for ik_mlp_forward_noalloc
    IK_HIP(hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking));
    m->m = ik::mlpCreate(m->s);
    std::vector<float> p;
    initParams(p, seed);
    IK_HIP(hipMemcpyAsync(m->m->P.p, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice, m->s));
    IK_HIP(hipStreamSynchronize(m->s));
  });
}

void ik_mlp_destroy(ik_mlp* m) {
  if (!m) return;
  ik::mlpDestroy(m->m);
  if (m->s) (void)hipStreamDestroy(m->s);
  delete m;
}

const char* ik_mlp_last_error(const ik_mlp* m) { return m ? m->err.c_str() : "null context"; }

int ik_mlp_params(ik_mlp* m, float* out, int where) {
  return mguard(m, [&] {
    IK_HIP(hipMemcpyAsync(out, m->m->P.p, ik::kParams * sizeof(float),
                          where == IK_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, m->s));
    IK_HIP(hipStreamSynchronize(m->s));
  });
}

int ik_mlp_load(ik_mlp* m, const float* p, int where) {
  return mguard(m, [&] {
    IK_HIP(hipMemcpyAsync(m->m->P.p, p, ik::kParams * sizeof(float),
                          where == IK_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, m->s));
    IK_HIP(hipStreamSynchronize(m->s));
    m->m->step = 0;
  });
}

int ik_mlp_forward(ik_mlp* m, const float* x, int64_t n, float* y, int where) {
  return mguard(m, [&] {
    ik::require(n > 0, IK_EINVAL, "empty batch");
    m->m->ensure(n);
    const float* xd = x;
    if (where != IK_DEVICE) {
      IK_HIP(hipMemcpyAsync(m->m->Xin.p, x, n * ik::kIn * sizeof(float), hipMemcpyHostToDevice, m->s));
      xd = m->m->Xin.p;
    }
    m->m->forward(xd, n);
    IK_HIP(hipMemcpyAsync(y, m->m->Y.p, n * ik::kOut * sizeof(float),
                          where == IK_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, m->s));
    IK_HIP(hipStreamSynchronize(m->s));
  });
}

int ik_mlp_forward_noalloc(ik_mlp* m, const float* x, int64_t n, float* y) {
  return mguard(m, [&] {
    ik::require(x != nullptr && y != nullptr, IK_EINVAL, "null argument");
    ik::require(n > 0 && n <= IK_MLP_NOALLOC_MAX_ROWS, IK_EINVAL,
                "ik_mlp_forward_noalloc: the batch must have 1 to IK_MLP_NOALLOC_MAX_ROWS (8192) rows");
    float* s0 = g_scratchPtr[m->device];  // resolved at ik_mlp_create; never the array's host-side name
    ik::require(s0 != nullptr, IK_EINVAL, "ik_mlp_forward_noalloc: scratch not resolved for this device");
    float* s1 = s0 + (int64_t)IK_MLP_NOALLOC_MAX_ROWS * ik::kHid;
    m->m->s = m->s;
    m->m->forwardInto(x, n, s0, s1, y);
    IK_HIP(hipStreamSynchronize(m->s));
  });
}

int ik_mlp_loss_grad(ik_mlp* m, const float* x, const float* t, int64_t n, double* loss, float* grad, int where) {
  return mguard(m, [&] {
    ik::require(n > 0, IK_EINVAL, "empty batch");
    m->m->ensure(n);
    const float *xd = x, *td = t;
    if (where != IK_DEVICE) {
      IK_HIP(hipMemcpyAsync(m->m->Xin.p, x, n * ik::kIn * sizeof(float), hipMemcpyHostToDevice, m->s));
      IK_HIP(hipMemcpyAsync(m->m->Tin.p, t, n * ik::kOut * sizeof(float), hipMemcpyHostToDevice, m->s));
      xd = m->m->Xin.p;
      td = m->m->Tin.p;
    }
    const double l = m->m->lossGrad(xd, td, n);
    if (loss) *loss = l;
    if (grad) {
      IK_HIP(hipMemcpyAsync(grad, m->m->G.p, ik::kParams * sizeof(float),
                            where == IK_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, m->s));
      IK_HIP(hipStreamSynchronize(m->s));
    }
  });
}

int ik_mlp_train_step(ik_mlp* m, const float* x, const float* t, int64_t n, float lr, double* loss, int where) {
  return mguard(m, [&] {
    ik::require(n > 0, IK_EINVAL, "empty batch");
    m->m->ensure(n);
    const float *xd = x, *td = t;
    if (where != IK_DEVICE) {
      IK_HIP(hipMemcpyAsync(m->m->Xin.p, x, n * ik::kIn * sizeof(float), hipMemcpyHostToDevice, m->s));
      IK_HIP(hipMemcpyAsync(m->m->Tin.p, t, n * ik::kOut * sizeof(float), hipMemcpyHostToDevice, m->s));
      xd = m->m->Xin.p;
      td = m->m->Tin.p;
    }
    const double l = m->m->lossGrad(xd, td, n);
    m->m->adam(lr, 0.9f, 0.999f, 1e-8f);
    IK_HIP(hipStreamSynchronize(m->s));
    if (loss) *loss = l;
  });
}

}  // extern "C"
