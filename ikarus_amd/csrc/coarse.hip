// coarse.hip — two-level preconditioner of the FP32 inner PCG: the 3x3 block-Jacobi smoother plus an additive
// coarse correction on aggregates of nodes spanned by the rigid-body modes,
//   z = D^-1 r + P (P^T K P)^-1 P^T r.
// No counterpart in the reference (its iterative solver is Eigen's Jacobi-preconditioned CG); the coarse space is
// the tentative prolongator of smoothed aggregation (Vanek, Mandel & Brezina, Computing 56 (1996) 179-196)
// without the smoothing step, with geometric box aggregates.
//
// Aggregates: boxes of a x a x a median element extents over the reference coordinates (the last box along each
// axis takes the remainder, so no aggregate is a thin layer); a is chosen from the node count so that there are
// about 729 aggregates (at 64^3 the fastest of 343, 512, 729 and 1000 on training load paths). Node i of
// aggregate A carries the six rigid-body modes about the aggregate's centroid, B(d_i) = [I | w -> w x d_i] with
// d_i = X_i - c_A, rows of constrained DOFs zero. Per aggregate the modes are
// orthonormalised over its nodes (T_A = V L^-1/2 from the eigenpairs of sum_i B_i^T B_i; modes with an eigenvalue
// below 1e-10 of the largest, e.g. rotations of a collinear aggregate, are dropped), so P_i = B(d_i) T_A and
// P^T K P is positive definite for positive definite K.
//
// Set-up (first two-level solve, again when a solve needs more than twice the inner iterations per decade of
// residual reduction of the first solve after the last set-up, plus 5): A_raw = sum over K blocks
// B_i^T K_ij B_j per aggregate pair on the device (FP64, fixed summation order), A = T^T A_raw T, Cholesky factorisation and inverse (rocSOLVER dpotrf / dpotri), and
// G = T A^-1 T^T stored dense in FP32. Per PCG iteration: s = B^T r per aggregate (one workgroup per aggregate),
// w = G s (one wavefront per row), z_i = D_i^-1 r_i + B(d_i) w_A (fused with the r.z partials).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>

#include "host_parallel.hh"
#include "ik_internal.hh"

namespace ik {

struct Coarse {
  bool ready = false;       // aggregates and T built
  bool valid = false;       // G holds the inverse of a current-enough coarse operator
  int64_t na = 0, nc = 0;   // aggregates, coarse DOFs (6 per aggregate)
  int64_t ld = 0;           // row stride of G and length of s: nc rounded up to 64 (padding zero)
  double rateAtBuild = -1;  // inner iterations per decade of the first solve after the last set-up
  int builds = 0;
  DevBuf<int32_t> agg, aggPtr, aggNodes, nbrPtr, nbr;
  DevBuf<float> d;          // [nn*3] node offsets from the aggregate centroid
  DevBuf<double> T;         // [na*36] row-major 6x6 per aggregate (columns of dropped modes zero)
  DevBuf<uint8_t> drop;     // [nc] dropped modes
  DevBuf<double> A;         // [nc*nc] coarse operator, then its inverse (column-major lower triangle)
  DevBuf<float> G;          // [nc*ld] T A^-1 T^T, row-major
  DevBuf<float> s, w;       // [ld], [nc]
  DevBuf<double> ws;        // [ld / 4] per-workgroup partials of w . s (coarse GEMV)
  DevBuf<int32_t> info;
  rocblas_handle blas = nullptr;
  ~Coarse() {
    if (blas) (void)rocblas_destroy_handle(blas);
  }
};

namespace {

constexpr int kTargetAggregates = 729;
constexpr int kMaxAggregates    = 1024;  // the update's 2 na partials fit the first 2 kReduceBlocks entries
// threads of the per-aggregate update workgroup (64^3 learned path: 14.3 us per launch with 256, 16.4 with 512,
// 51.9 with 1024)
constexpr int kUpdateThreads = 256;

__device__ inline double waveSumD(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline float waveSumF(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// B(d)^T v for the masked modes: [v; d x v] with v's constrained components zeroed
template <typename T>
__device__ inline void modesT(const T v[3], const float d[3], T out[6]) {
  out[0] = v[0];
  out[1] = v[1];
  out[2] = v[2];
  out[3] = (T)d[1] * v[2] - (T)d[2] * v[1];
  out[4] = (T)d[2] * v[0] - (T)d[0] * v[2];
  out[5] = (T)d[0] * v[1] - (T)d[1] * v[0];
}
// B(d) w: the translation plus the rotation w[3..5] x d
template <typename T>
__device__ inline void modes(const T w[6], const float d[3], T out[3]) {
  out[0] = w[0] + w[4] * (T)d[2] - w[5] * (T)d[1];
  out[1] = w[1] + w[5] * (T)d[0] - w[3] * (T)d[2];
  out[2] = w[2] + w[3] * (T)d[1] - w[4] * (T)d[0];
}

// A_raw(A, B) = sum over nodes i of A and blocks (i, j) with j in aggregate B of B_i^T K_ij B_j. One workgroup per
// aggregate A, one pass per neighbouring aggregate B; each thread sums its nodes' blocks, then the workgroup
// reduces the 36 sums in a fixed order. K in sliced-ELL slots.
template <typename TV>
__global__ __launch_bounds__(256) void coarseAssemble(const TV* __restrict__ K, const int32_t* __restrict__ rowptr,
                                                      const int32_t* __restrict__ sellStart,
                                                      const int32_t* __restrict__ scol,
                                                      const int32_t* __restrict__ agg,
                                                      const int32_t* __restrict__ aggPtr,
                                                      const int32_t* __restrict__ aggNodes,
                                                      const int32_t* __restrict__ nbrPtr,
                                                      const int32_t* __restrict__ nbr, const float* __restrict__ dd,
                                                      const uint8_t* __restrict__ fixed, int64_t nc,
                                                      double* __restrict__ Araw) {
  __shared__ double red[4][36];
  const int a = blockIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int nb = nbrPtr[a]; nb < nbrPtr[a + 1]; ++nb) {
    const int b = nbr[nb];
    double acc[36];
#pragma unroll
    for (int m = 0; m < 36; ++m) acc[m] = 0.0;
    for (int t = aggPtr[a] + threadIdx.x; t < aggPtr[a + 1]; t += 256) {
      const int32_t i = aggNodes[t];
      const float di[3] = {dd[3 * i], dd[3 * i + 1], dd[3 * i + 2]};
      const bool fi[3] = {fixed[3 * i] != 0, fixed[3 * i + 1] != 0, fixed[3 * i + 2] != 0};
      const int len      = rowptr[i + 1] - rowptr[i];
      const int64_t base = (int64_t)sellStart[i >> 6] + (i & 63);
      for (int j = 0; j < len; ++j) {
        const int64_t slot = base + (int64_t)j * kSellRows;
        const int32_t c    = scol[slot];
        if (agg[c] != b) continue;
        const float dc[3] = {dd[3 * c], dd[3 * c + 1], dd[3 * c + 2]};
        // M = K_ij B_c (3x6): column l of B_c is the displacement of mode l at node c (masked)
        double M[3][6];
        for (int l = 0; l < 6; ++l) {
          double e[6] = {0, 0, 0, 0, 0, 0}, col[3];
          e[l] = 1.0;
          modes<double>(e, dc, col);
          for (int q = 0; q < 3; ++q)
            if (fixed[3 * c + q]) col[q] = 0.0;
          for (int k = 0; k < 3; ++k)
            M[k][l] = (double)K[slot * 9 + 3 * k] * col[0] + (double)K[slot * 9 + 3 * k + 1] * col[1] +
                      (double)K[slot * 9 + 3 * k + 2] * col[2];
        }
        // acc += B_i^T M
        for (int l = 0; l < 6; ++l) {
          double v[3] = {fi[0] ? 0.0 : M[0][l], fi[1] ? 0.0 : M[1][l], fi[2] ? 0.0 : M[2][l]}, o[6];
          modesT<double>(v, di, o);
#pragma unroll
          for (int m = 0; m < 6; ++m) acc[m * 6 + l] += o[m];
        }
      }
    }
#pragma unroll
    for (int m = 0; m < 36; ++m) {
      const double v = waveSumD(acc[m]);
      if (lane == 0) red[wave][m] = v;
    }
    __syncthreads();
    if (threadIdx.x < 36) {
      const int m = threadIdx.x;
      const double v = ((red[0][m] + red[1][m]) + red[2][m]) + red[3][m];
      Araw[(int64_t)(6 * a + m / 6) * nc + 6 * b + m % 6] = v;
    }
    __syncthreads();
  }
}

// A(a, b) = T_a^T A_raw(a, b) T_b per 6x6 block; a unit diagonal for dropped modes (their rows and columns are
// zero otherwise). Thread per (a, b, entry).
__global__ void coarseCongruence(const double* __restrict__ Araw, const double* __restrict__ T,
                                 const uint8_t* __restrict__ drop, int64_t na, double* __restrict__ A) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t nc  = 6 * na;
  if (idx >= nc * nc) return;
  const int64_t r = idx / nc, c = idx - r * nc;
  const int64_t a = r / 6, b = c / 6;
  const int m = (int)(r - 6 * a), l = (int)(c - 6 * b);
  double v = 0.0;
  for (int p = 0; p < 6; ++p) {
    double t = 0.0;
    for (int q = 0; q < 6; ++q) t += Araw[(6 * a + p) * nc + 6 * b + q] * T[b * 36 + q * 6 + l];
    v += T[a * 36 + p * 6 + m] * t;
  }
  if (r == c && drop[r]) v = 1.0;
  A[idx] = v;  // symmetric: the row-major and column-major views agree
}

// G = T A^-1 T^T in FP32 (row stride ld) from the inverse's lower triangle (column-major, entry (i, j) for i >= j
// at j nc + i)
__global__ void coarseInverseOut(const double* __restrict__ Ainv, const double* __restrict__ T, int64_t na,
                                 int64_t ld, float* __restrict__ G) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t nc  = 6 * na;
  if (idx >= nc * nc) return;
  const int64_t r = idx / nc, c = idx - r * nc;
  const int64_t a = r / 6, b = c / 6;
  const int m = (int)(r - 6 * a), l = (int)(c - 6 * b);
  double v = 0.0;
  for (int p = 0; p < 6; ++p) {
    double t = 0.0;
    for (int q = 0; q < 6; ++q) {
      const int64_t i = 6 * a + p, j = 6 * b + q;
      const double e  = i >= j ? Ainv[j * nc + i] : Ainv[i * nc + j];
      t += e * T[b * 36 + l * 6 + q];
    }
    v += T[a * 36 + m * 6 + p] * t;
  }
  G[r * ld + c] = (float)v;
}

// s_a = sum over the nodes of aggregate a of B(d_i)^T r_i (constrained components zeroed): one workgroup per
// aggregate, fixed-order reduction
__global__ __launch_bounds__(256) void coarseRestrict(const float* __restrict__ r, const uint8_t* __restrict__ fixed,
                                                      const float* __restrict__ dd,
                                                      const int32_t* __restrict__ aggPtr,
                                                      const int32_t* __restrict__ aggNodes, float* __restrict__ s,
                                                      const PcgScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ float red[4][6];
  const int a = blockIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float acc[6] = {0, 0, 0, 0, 0, 0};
  for (int t = aggPtr[a] + threadIdx.x; t < aggPtr[a + 1]; t += 256) {
    const int32_t i = aggNodes[t];
    float v[3], o[6];
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = fixed[3 * i + k] ? 0.0f : r[3 * i + k];
    const float di[3] = {dd[3 * i], dd[3 * i + 1], dd[3 * i + 2]};
    modesT<float>(v, di, o);
#pragma unroll
    for (int m = 0; m < 6; ++m) acc[m] += o[m];
  }
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    const float v = waveSumF(acc[m]);
    if (lane == 0) red[wave][m] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int m = threadIdx.x;
    s[6 * a + m] = ((red[0][m] + red[1][m]) + red[2][m]) + red[3][m];
  }
}

// w = G s: one wavefront per row, 4 rows per workgroup, 16-byte loads (rows padded to ld = a multiple of 64 floats)
__global__ __launch_bounds__(256) void coarseGemv(const float4* __restrict__ G, const float4* __restrict__ s,
                                                  float* __restrict__ w, int64_t nc, int64_t ld4,
                                                  const PcgScalars* __restrict__ sc) {
  if (sc->done) return;
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= nc) return;
  const float4* g = G + row * ld4;
  float acc = 0.0f;
  for (int64_t c = lane; c < ld4; c += 64) {
    const float4 a = g[c], b = s[c];
    acc += a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
  }
  acc = waveSumF(acc);
  if (lane == 0) w[row] = acc;
}

// z_i = D_i^-1 r_i + B(d_i) w_a (constrained components of the coarse part zeroed), optionally p = z; partials
// of r.z into partial[grid + block]
__global__ __launch_bounds__(256) void coarseZ(const float* __restrict__ r, const float* __restrict__ binv,
                                               const uint8_t* __restrict__ fixed, const float* __restrict__ dd,
                                               const int32_t* __restrict__ agg, const float* __restrict__ w,
                                               float* __restrict__ z, float* __restrict__ p, int64_t nn,
                                               double* __restrict__ partial, const PcgScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ double red[4];
  double rz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nn; i += (int64_t)gridDim.x * 256) {
    const float rl[3] = {r[3 * i], r[3 * i + 1], r[3 * i + 2]};
    const float di[3] = {dd[3 * i], dd[3 * i + 1], dd[3 * i + 2]};
    const int32_t a   = agg[i];
    float wa[6], cz[3];
#pragma unroll
    for (int m = 0; m < 6; ++m) wa[m] = w[6 * a + m];
    modes<float>(wa, di, cz);
    const float* B = binv + i * 9;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float zk = B[3 * k] * rl[0] + B[3 * k + 1] * rl[1] + B[3 * k + 2] * rl[2];
      if (!fixed[3 * i + k]) zk += cz[k];
      z[3 * i + k] = zk;
      if (p) p[3 * i + k] = zk;
      rz += (double)rl[k] * (double)zk;
    }
  }
  rz = waveSumD(rz);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = rz;
  __syncthreads();
  if (threadIdx.x == 0) partial[gridDim.x + blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// N workgroup sums (W wavefronts) at once, in a fixed order; the results are valid in every thread. One barrier
// pair for all N (the wave reductions run side by side).
template <int W, int N>
__device__ inline void groupSums(double (&v)[N], double (&red)[W][N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = waveSumD(v[i]);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int i = 0; i < N; ++i) red[threadIdx.x >> 6][i] = v[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double t = red[0][i];
#pragma unroll
    for (int w = 1; w < W; ++w) t += red[w][i];
    v[i] = t;
  }
  __syncthreads();
}

// PCG iteration of the two-level form, part 1 (one workgroup per aggregate, over its nodes): alpha = r.z / p.q
// (every workgroup folds the npq p.q partials in the same order; workgroup 0 records the scalars),
// x += alpha p, r -= alpha q, and for the new r: the r.r and r.D^-1 r partials (partial[a], partial[na + a]) and
// the restriction s_a = sum B(d_i)^T r_i (constrained components zeroed). With w = G s, r.z = r.D^-1 r + w.s,
// since r . B(d_i) w_a summed over the nodes is w_a . s_a.
template <int BS>
__global__ __launch_bounds__(BS) void coarseUpdate(float* __restrict__ x, float* __restrict__ r,
                                                    const float* __restrict__ p, const float* __restrict__ q,
                                                    const float* __restrict__ binv, const uint8_t* __restrict__ fixed,
                                                    const float* __restrict__ dd, const int32_t* __restrict__ aggPtr,
                                                    const int32_t* __restrict__ aggNodes, int na, float* __restrict__ s,
                                                    const double* __restrict__ pq, int npq,
                                                    double* __restrict__ partial, PcgScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ double red1[BS / 64][1];
  __shared__ double red8[BS / 64][8];
  const int a = blockIdx.x;
  double pqv[1] = {0.0};
  for (int i = threadIdx.x; i < npq; i += BS) pqv[0] += pq[i];
  groupSums<BS / 64, 1>(pqv, red1);
  const double rz = sc->rz_new;  // r.z of the current iterate (set by the previous p update or the init fold)
  const double alpha = rz / pqv[0];
  if (a == 0 && threadIdx.x == 0) {
    sc->rz    = rz;
    sc->pq    = pqv[0];
    sc->alpha = alpha;
  }
  const float al = (float)alpha;
  double rr = 0.0, rzb = 0.0;
  float sa[6] = {0, 0, 0, 0, 0, 0};
  for (int t = aggPtr[a] + threadIdx.x; t < aggPtr[a + 1]; t += BS) {
    const int64_t i = aggNodes[t];
    float rl[3], v[3], o[6];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int64_t j = 3 * i + k;
      x[j] += al * p[j];
      rl[k] = r[j] - al * q[j];
      r[j]  = rl[k];
      rr += (double)rl[k] * (double)rl[k];
      v[k] = fixed[j] ? 0.0f : rl[k];
    }
    const float* B = binv + i * 9;
#pragma unroll
    for (int k = 0; k < 3; ++k)
      rzb += (double)rl[k] * (double)(B[3 * k] * rl[0] + B[3 * k + 1] * rl[1] + B[3 * k + 2] * rl[2]);
    const float di[3] = {dd[3 * i], dd[3 * i + 1], dd[3 * i + 2]};
    modesT<float>(v, di, o);
#pragma unroll
    for (int m = 0; m < 6; ++m) sa[m] += o[m];
  }
  double v[8] = {rr, rzb, sa[0], sa[1], sa[2], sa[3], sa[4], sa[5]};
  groupSums<BS / 64, 8>(v, red8);
  if (threadIdx.x == 0) {
    partial[a]      = v[0];
    partial[na + a] = v[1];
#pragma unroll
    for (int m = 0; m < 6; ++m) s[6 * a + m] = (float)v[2 + m];
  }
}

// part 2: w = G s (as coarseGemv) and per workgroup the partial of w . s
__global__ __launch_bounds__(256) void coarseGemvDot(const float4* __restrict__ G, const float4* __restrict__ s4,
                                                     const float* __restrict__ s, float* __restrict__ w, int64_t nc,
                                                     int64_t ld4, double* __restrict__ ws,
                                                     const PcgScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ double red[4];
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  float acc = 0.0f;
  if (row < nc) {
    const float4* g = G + row * ld4;
    for (int64_t c = lane; c < ld4; c += 64) {
      // G (77 MB at 64^3) is read once per iteration and evicted by the matrix before the next: loaded
      // non-temporally it leaves the last-level cache to the SpMV's matrix lines (two-level 64^3 path 2-5 % faster)
      typedef float f4v __attribute__((ext_vector_type(4)));
      const f4v a   = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(g + c));
      const float4 b = s4[c];
      acc += a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
    }
  }
  acc = waveSumF(acc);
  if (lane == 0) {
    if (row < nc) w[row] = acc;
    red[threadIdx.x >> 6] = row < nc ? (double)acc * (double)s[row] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) ws[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// part 3: every workgroup folds r.r, r.D^-1 r (2 na partials) and w.s (nws partials) in the same order; the
// convergence test and beta = r.z / r.z_prev, then p = z + beta p with z_i = D_i^-1 r_i + B(d_i) w_a
// (constrained components of the coarse part zeroed) formed on the fly
__global__ __launch_bounds__(256) void coarseBetaP(float* __restrict__ p, const float* __restrict__ r,
                                                   const float* __restrict__ binv, const uint8_t* __restrict__ fixed,
                                                   const float* __restrict__ dd, const int32_t* __restrict__ agg,
                                                   const float* __restrict__ w, int64_t nn,
                                                   const double* __restrict__ partial, int na,
                                                   const double* __restrict__ ws, int nws,
                                                   PcgScalars* __restrict__ sc, int maxIter) {
  if (sc->done) return;
  __shared__ double red3[4][3];
  double f[3] = {0.0, 0.0, 0.0};  // r.r, r.D^-1 r, w.s
  for (int i = threadIdx.x; i < na; i += 256) {
    f[0] += partial[i];
    f[1] += partial[na + i];
  }
  for (int i = threadIdx.x; i < nws; i += 256) f[2] += ws[i];
  groupSums<4, 3>(f, red3);
  const double rr = f[0], rz = f[1] + f[2];
  const bool conv = rr < sc->threshold;
  const double bt = rz / sc->rz;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->rr = rr;
    if (conv) {
      sc->done = 1;
    } else {
      sc->beta   = bt;
      sc->rz_new = rz;
      sc->iters += 1;
      if (sc->iters >= maxIter) sc->done = 2;
    }
  }
  if (conv) return;
  const float beta = (float)bt;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nn; i += (int64_t)gridDim.x * 256) {
    const float rl[3] = {r[3 * i], r[3 * i + 1], r[3 * i + 2]};
    const float di[3] = {dd[3 * i], dd[3 * i + 1], dd[3 * i + 2]};
    const int32_t a   = agg[i];
    float wa[6], cz[3];
#pragma unroll
    for (int m = 0; m < 6; ++m) wa[m] = w[6 * a + m];
    modes<float>(wa, di, cz);
    const float* B = binv + i * 9;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float zk = B[3 * k] * rl[0] + B[3 * k + 1] * rl[1] + B[3 * k + 2] * rl[2];
      if (!fixed[3 * i + k]) zk += cz[k];
      p[3 * i + k] = zk + beta * p[3 * i + k];
    }
  }
}

// cyclic Jacobi eigen-decomposition of a symmetric 6x6 matrix (host): a becomes diagonal, v the eigenvectors
// (columns)
void jacobiEigen6(double a[6][6], double v[6][6]) {
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) v[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) (i == j ? tot : off) += a[i][j] * a[i][j];
    if (off <= 1e-30 * (tot + off)) break;
    for (int p = 0; p < 5; ++p)
      for (int q = p + 1; q < 6; ++q) {
        if (a[p][q] == 0.0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double t     = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 6; ++k) {  // columns p, q
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 6; ++k) {  // rows p, q
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 6; ++k) {
          const double vkp = v[k][p], vkq = v[k][q];
          v[k][p] = c * vkp - s * vkq;
          v[k][q] = s * vkp + c * vkq;
        }
      }
  }
}

// aggregates, node offsets, orthonormalising transforms and aggregate adjacency (host, once per handle)
void buildAggregates(ik_handle* h, Coarse& cs) {
  const int64_t nn = h->nn;
  const double* X = h->hX.data();
  // boxes of a x a x a cells of the mesh lattice (node bounds, median element extent per axis)
  const double* lo = h->latLo;
  const double* hi = h->latHi;
  const double* step = h->latStep;
  auto boxes = [&](int a, int64_t nb[3]) {
    for (int d = 0; d < 3; ++d)
      nb[d] = std::max<int64_t>(1, (int64_t)std::floor((hi[d] - lo[d]) / (a * step[d]) + 1e-9));
    return nb[0] * nb[1] * nb[2];
  };
  int a = std::max(2, (int)std::lround(std::cbrt((double)nn / kTargetAggregates)));
  int64_t nb[3];
  while (boxes(a, nb) > kMaxAggregates) ++a;
  // box of each node, then compact aggregate ids in box order
  HostVec<int64_t> key(nn);
  parallelChunks(nn, [&](int64_t n0, int64_t n1, int64_t) {
    for (int64_t n = n0; n < n1; ++n) {
      int64_t k[3];
      for (int d = 0; d < 3; ++d) {
        const int64_t c = (int64_t)std::floor((X[n * 3 + d] - lo[d]) / (a * step[d]) + 1e-9);
        k[d]            = std::min<int64_t>(nb[d] - 1, std::max<int64_t>(0, c));
      }
      key[n] = (k[2] * nb[1] + k[1]) * nb[0] + k[0];
    }
  });
  std::vector<int32_t> boxId(nb[0] * nb[1] * nb[2], -1);
  HostVec<int32_t> agg(nn);
  int32_t na = 0;
  for (int64_t n = 0; n < nn; ++n)
    if (boxId[key[n]] < 0) boxId[key[n]] = 0;
  for (auto& b : boxId)
    if (b == 0) b = na++;
  parallelChunks(nn, [&](int64_t n0, int64_t n1, int64_t) {
    for (int64_t n = n0; n < n1; ++n) agg[n] = boxId[key[n]];
  });
  // nodes of each aggregate, ascending
  std::vector<int32_t> ptr(na + 1, 0);
  HostVec<int32_t> nodes(nn);
  for (int64_t n = 0; n < nn; ++n) ptr[agg[n] + 1]++;
  for (int32_t k = 0; k < na; ++k) ptr[k + 1] += ptr[k];
  {
    std::vector<int32_t> pos(ptr.begin(), ptr.end() - 1);
    for (int64_t n = 0; n < nn; ++n) nodes[pos[agg[n]]++] = (int32_t)n;
  }
  // centroids (node sums in ascending node order) and offsets
  std::vector<double> cen((size_t)na * 3, 0.0);
  parallelChunks(na, [&](int64_t k0, int64_t k1, int64_t) {
    for (int64_t k = k0; k < k1; ++k) {
      for (int32_t t = ptr[k]; t < ptr[k + 1]; ++t)
        for (int d = 0; d < 3; ++d) cen[k * 3 + d] += X[(int64_t)nodes[t] * 3 + d];
      for (int d = 0; d < 3; ++d) cen[k * 3 + d] /= (ptr[k + 1] - ptr[k]);
    }
  }, 16);
  HostVec<float> dd(nn * 3);
  parallelChunks(nn, [&](int64_t n0, int64_t n1, int64_t) {
    for (int64_t n = n0; n < n1; ++n)
      for (int d = 0; d < 3; ++d) dd[n * 3 + d] = (float)(X[n * 3 + d] - cen[agg[n] * 3 + d]);
  });
  // T_a from the Gram matrix of the masked modes (evaluated with the same float offsets the device uses)
  std::vector<double> T((size_t)na * 36, 0.0);
  std::vector<uint8_t> drop((size_t)na * 6, 0);
  parallelChunks(na, [&](int64_t k0, int64_t k1, int64_t) {
    for (int64_t k = k0; k < k1; ++k) {
      double g[6][6] = {}, v[6][6];
      for (int32_t t = ptr[k]; t < ptr[k + 1]; ++t) {
        const int32_t n = nodes[t];
        const double dx = dd[n * 3], dy = dd[n * 3 + 1], dz = dd[n * 3 + 2];
        const double Bm[3][6] = {{1, 0, 0, 0, dz, -dy}, {0, 1, 0, -dz, 0, dx}, {0, 0, 1, dy, -dx, 0}};
        for (int r = 0; r < 3; ++r) {
          if (h->hfixed[3 * n + r]) continue;
          for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) g[i][j] += Bm[r][i] * Bm[r][j];
        }
      }
      double g0[6][6];
      std::copy(&g[0][0], &g[0][0] + 36, &g0[0][0]);
      jacobiEigen6(g, v);
      double lmax = 0.0;
      for (int i = 0; i < 6; ++i) lmax = std::max(lmax, g[i][i]);
      for (int j = 0; j < 6; ++j) {
        const double l = g[j][j];
        if (!(lmax > 0.0) || !(l > 1e-10 * lmax)) {
          drop[k * 6 + j] = 1;
          continue;
        }
        for (int i = 0; i < 6; ++i) T[k * 36 + i * 6 + j] = v[i][j] / std::sqrt(l);
      }
      // the kept modes are orthonormal over the aggregate: T^T G T = I on them
      for (int p = 0; p < 6; ++p)
        for (int q = 0; q < 6; ++q) {
          if (drop[k * 6 + p] || drop[k * 6 + q]) continue;
          double e = 0.0;
          for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) e += T[k * 36 + i * 6 + p] * g0[i][j] * T[k * 36 + j * 6 + q];
          require(std::fabs(e - (p == q ? 1.0 : 0.0)) < 1e-6, IK_EINVAL,
                  "internal error: aggregate modes not orthonormalised");
        }
    }
  }, 16);
  // aggregate adjacency from the block pattern: per chunk of aggregates, each aggregate's sorted neighbours,
  // placed by the counts afterwards
  std::vector<int32_t> nptr(na + 1, 0), nlist;
  {
    std::vector<std::vector<int32_t>> part(chunkCount(na, 16));
    std::vector<int32_t> cnt(na);
    parallelChunks(na, [&](int64_t k0, int64_t k1, int64_t c) {
      std::vector<int32_t> mark(na, -1), tmp;
      for (int64_t k = k0; k < k1; ++k) {
        tmp.clear();
        for (int32_t t = ptr[k]; t < ptr[k + 1]; ++t) {
          const int32_t n = nodes[t];
          for (int32_t b = h->hrowptr[n]; b < h->hrowptr[n + 1]; ++b) {
            const int32_t o = agg[h->hcolidx[b]];
            if (mark[o] != (int32_t)k) {
              mark[o] = (int32_t)k;
              tmp.push_back(o);
            }
          }
        }
        std::sort(tmp.begin(), tmp.end());
        part[c].insert(part[c].end(), tmp.begin(), tmp.end());
        cnt[k] = (int32_t)tmp.size();
      }
    }, 16);
    for (int32_t k = 0; k < na; ++k) nptr[k + 1] = nptr[k] + cnt[k];
    nlist.reserve(nptr[na]);
    for (const auto& v : part) nlist.insert(nlist.end(), v.begin(), v.end());
  }
  cs.na = na;
  cs.nc = 6 * (int64_t)na;
  cs.ld = (cs.nc + 63) / 64 * 64;
  cs.agg.upload(agg.data(), agg.size(), h->stream);
  cs.aggPtr.upload(ptr.data(), ptr.size(), h->stream);
  cs.aggNodes.upload(nodes.data(), nodes.size(), h->stream);
  cs.nbrPtr.upload(nptr.data(), nptr.size(), h->stream);
  cs.nbr.upload(nlist.data(), nlist.size(), h->stream);
  cs.d.upload(dd.data(), dd.size(), h->stream);
  cs.T.upload(T.data(), T.size(), h->stream);
  cs.drop.upload(drop.data(), drop.size(), h->stream);
  cs.s.ensure(cs.ld);
  launchZeroF(cs.s.p, cs.ld, h->stream);
  cs.w.ensure(cs.nc);
  cs.ws.ensure(cs.ld / 4);
  cs.info.ensure(2);
  cs.ready = true;
}

}  // namespace

void coarseDestroy(Coarse* c) { delete c; }

int64_t coarseDofs(const ik_handle* h) { return h->coarse ? h->coarse->nc : 0; }

// (Re)builds the coarse inverse from the current FP32 tangent when there is none yet or when the last solve
// needed more than twice the inner iterations per decade of residual reduction of the first solve after the last
// build, plus 5 (the tangent has drifted from the one the coarse operator came from).
void coarsePrepare(ik_handle* h) {
  if (!h->coarse) h->coarse = new Coarse();
  Coarse& cs = *h->coarse;
  if (!cs.ready) buildAggregates(h, cs);
  if (cs.valid) return;
  finishErrCheck(h, true);  // a tangent with an invalid element reports that, not a failed coarse factorisation
  const int64_t nc = cs.nc;
  DevBuf<double> raw;
  raw.ensure((size_t)(nc * nc));
  cs.A.ensure((size_t)(nc * nc));
  if (!cs.G.p) {
    cs.G.ensure((size_t)(nc * cs.ld));
    launchZeroF(cs.G.p, nc * cs.ld, h->stream);  // the padding columns stay zero
  }
  launchZero(raw.p, nc * nc, h->stream);
  hipLaunchKernelGGL(coarseAssemble<float>, dim3((unsigned)cs.na), dim3(256), 0, h->stream, h->K32.p, h->rowptr.p,
                     h->sellStart.p, h->scol.p, cs.agg.p, cs.aggPtr.p, cs.aggNodes.p, cs.nbrPtr.p, cs.nbr.p, cs.d.p,
                     h->fixed.p, nc, raw.p);
  const unsigned g2 = (unsigned)((nc * nc + 255) / 256);
  hipLaunchKernelGGL(coarseCongruence, dim3(g2), dim3(256), 0, h->stream, raw.p, cs.T.p, cs.drop.p, cs.na, cs.A.p);
  if (!cs.blas) {
    require(rocblas_create_handle(&cs.blas) == rocblas_status_success, IK_EHIP, "rocblas_create_handle failed");
  }
  require(rocblas_set_stream(cs.blas, h->stream) == rocblas_status_success, IK_EHIP, "rocblas_set_stream failed");
  require(rocsolver_dpotrf(cs.blas, rocblas_fill_lower, (rocblas_int)nc, cs.A.p, (rocblas_int)nc, cs.info.p) ==
              rocblas_status_success,
          IK_EHIP, "rocsolver_dpotrf failed");
  require(rocsolver_dpotri(cs.blas, rocblas_fill_lower, (rocblas_int)nc, cs.A.p, (rocblas_int)nc, cs.info.p + 1) ==
              rocblas_status_success,
          IK_EHIP, "rocsolver_dpotri failed");
  int32_t info[2] = {0, 0};
  IK_HIP(hipMemcpyAsync(info, cs.info.p, sizeof(info), hipMemcpyDeviceToHost, h->stream));
  IK_HIP(hipStreamSynchronize(h->stream));
  require(info[0] == 0 && info[1] == 0, IK_ESINGULAR,
          "coarse operator of the two-level preconditioner is not positive definite (indefinite tangent)");
  hipLaunchKernelGGL(coarseInverseOut, dim3(g2), dim3(256), 0, h->stream, cs.A.p, cs.T.p, cs.na, cs.ld, cs.G.p);
  IK_HIP(hipStreamSynchronize(h->stream));  // raw is released on return
  cs.valid       = true;
  cs.rateAtBuild = -1;
  cs.builds++;
}

void coarseRecord(ik_handle* h, int iterations, double relres) {
  Coarse& cs = *h->coarse;
  const double decades = std::max(0.5, -std::log10(std::max(relres, 1e-12)));
  const double rate    = iterations / decades;
  if (cs.rateAtBuild < 0)
    cs.rateAtBuild = rate;
  else if (rate > 2.0 * cs.rateAtBuild + 5.0)
    cs.valid = false;
}

int coarseBuilds(const ik_handle* h) { return h->coarse ? h->coarse->builds : 0; }

void coarseApply(ik_handle* h, const float* r, float* z, float* p, const float* binv, int gn, double* partial) {
  Coarse& cs = *h->coarse;
  hipLaunchKernelGGL(coarseRestrict, dim3((unsigned)cs.na), dim3(256), 0, h->stream, r, h->fixed.p, cs.d.p,
                     cs.aggPtr.p, cs.aggNodes.p, cs.s.p, h->sc.p);
  hipLaunchKernelGGL(coarseGemv, dim3((unsigned)((cs.nc + 3) / 4)), dim3(256), 0, h->stream, (const float4*)cs.G.p,
                     (const float4*)cs.s.p, cs.w.p, cs.nc, cs.ld / 4, h->sc.p);
  hipLaunchKernelGGL(coarseZ, dim3(gn), dim3(256), 0, h->stream, r, binv, h->fixed.p, cs.d.p, cs.agg.p, cs.w.p, z, p,
                     h->nn, partial, h->sc.p);
}

// the p.q fold, update and restriction, coarse GEMV and p update of one two-level PCG iteration (grid gn for the p
// update; the update's partials occupy partial[0, 2 na))
void coarseIteration(ik_handle* h, const double* pq, int npq, float* x, float* r, float* p, const float* q,
                     const float* binv, int gn, double* partial, int maxIter) {
  Coarse& cs = *h->coarse;
  hipLaunchKernelGGL(coarseUpdate<kUpdateThreads>, dim3((unsigned)cs.na), dim3(kUpdateThreads), 0, h->stream, x, r,
                     p, q, binv, h->fixed.p,
                     cs.d.p, cs.aggPtr.p, cs.aggNodes.p, (int)cs.na, cs.s.p, pq, npq, partial, h->sc.p);
  const unsigned gg = (unsigned)((cs.nc + 3) / 4);
  hipLaunchKernelGGL(coarseGemvDot, dim3(gg), dim3(256), 0, h->stream, (const float4*)cs.G.p, (const float4*)cs.s.p,
                     cs.s.p, cs.w.p, cs.nc, cs.ld / 4, cs.ws.p, h->sc.p);
  hipLaunchKernelGGL(coarseBetaP, dim3(gn), dim3(256), 0, h->stream, p, r, binv, h->fixed.p, cs.d.p, cs.agg.p, cs.w.p,
                     h->nn, partial, (int)cs.na, cs.ws.p, (int)gg, h->sc.p, maxIter);
}

}  // namespace ik
