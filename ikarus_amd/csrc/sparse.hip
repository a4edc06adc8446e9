// sparse.hip — sliced-ELL 3x3-block SpMV, reductions and the preconditioned CG solvers.
//
// FP64 Jacobi-PCG follows Eigen's ConjugateGradient with DiagonalPreconditioner, the reference's
// si_ConjugateGradient (ikarus/solver/linearsolver/linearsolver.cpp:23-24): x0 = 0,
// stop when ||r||^2 < tol^2 ||b||^2, iteration count as Eigen reports it.
// The mixed form runs the same recurrence in FP32 on an FP32 tangent with a 3x3 block-Jacobi
// preconditioner, wrapped in FP64 iterative refinement (r = b - K x accumulated in FP64).
//
// All PCG scalars live on the device; each kernel reads a `done` flag so the host can enqueue
// iterations in chunks and poll the flag asynchronously: the iterate after convergence is exactly the
// one Eigen's loop returns.
#include "ik_internal.hh"

namespace ik {
namespace {

constexpr int SPMV_BS = 256;  // threads per block

// Block rows per wavefront of the SpMV lane layouts: FP64 values 3 lanes per row (21 rows, lane 63 idles),
// FP32 values one lane per row (64 rows).
template <typename TV>
constexpr int spmvRowsPerWave() {
  return sizeof(TV) == 4 ? 64 : 21;
}
template <typename TV>
constexpr int spmvRowsPerBlock() {
  return (SPMV_BS / 64) * spmvRowsPerWave<TV>();
}
template <typename TV>
inline unsigned spmvGrid(int64_t nrows) {
  return (unsigned)((nrows + spmvRowsPerBlock<TV>() - 1) / spmvRowsPerBlock<TV>());
}

__device__ inline double waveSum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// block-wide sum; result valid in thread 0
template <int BS>
__device__ inline double blockSum(double v) {
  __shared__ double red[BS / 64];
  v = waveSum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < BS / 64; ++i) s += red[i];
  __syncthreads();
  return s;
}

// y = K x over the sliced-ELL block storage (slice of kSellRows = 64 block rows, block j of row r of slice s
// at slot sellStart[s] + 64 j + r, 9 values per slot); optional fused partial dot d = sum_i x_i y_i per
// block. Two lane layouts:
//  FP64 values: lane (r, k), row r = lane / 3 (21 rows per wavefront), output component k: for a fixed j the
//   wavefront's 21 rows read 21 adjacent 72-byte blocks (1,512 contiguous bytes), each lane its block row.
//   Each row runs over its own blocks in column order, so the result equals a CSR SpMV's bit for bit.
//  FP32 values: one lane per block row, the 3 outputs of the row per lane, a wavefront = one slice: for a
//   fixed j the 64 lanes read 64 adjacent 36-byte blocks (2304 contiguous bytes). Each lane runs to the
//   slice's longest row; padding slots hold zeros with the row's own column.
// 64^3 pattern, alone: FP32 48.6 us against 62.3 us for 3 lanes per row over CSR (21 rows per wavefront);
// FP64 108.4 us against 119.3 us for the former layout with one lane per block entry (7 rows per wavefront).
// FP64 issues the loads of SPMV_UNROLL blocks (3 x 24 bytes and 3 column indices per lane) before summing their
// products: 3 beat 2, 4, 5, 6 and 9 (111.3, 111.9, 114.8, 116.5 and 121.1 us).
constexpr int SPMV_UNROLL = 3;
static_assert(kSellRows == 64, "one slice per wavefront");

// a[k] = (K x)[3 row + k], k = 0..2, for one lane per row: the slice's w slots of row `row`
template <typename TV, typename TX, typename TA>
__device__ inline void rowSell3(const int32_t* __restrict__ sellStart, const int32_t* __restrict__ scol,
                                const TV* __restrict__ val, const TX* __restrict__ x, int64_t row, TA a[3]) {
  const int64_t s = row >> 6;
  const int r     = (int)(row & 63);
  const int32_t s0 = sellStart[s];
  const int w      = (sellStart[s + 1] - s0) >> 6;
  const TV* V      = val + ((int64_t)s0 + r) * 9;
  const int32_t* C = scol + s0 + r;
  TA a0 = TA(0), a1 = TA(0), a2 = TA(0);
  // 3 slots per unrolled step: 53.2 us at 64^3 against 56.3, 60.9 and 68.4 us for 2, 4 and 6
#pragma unroll 3
  for (int j = 0; j < w; ++j) {
    const int64_t c = (int64_t)C[j * kSellRows] * 3;
    const TA x0 = (TA)x[c], x1 = (TA)x[c + 1], x2 = (TA)x[c + 2];
    const TV* v = V + (int64_t)j * (kSellRows * 9);
    a0 += (TA)v[0] * x0 + (TA)v[1] * x1 + (TA)v[2] * x2;
    a1 += (TA)v[3] * x0 + (TA)v[4] * x1 + (TA)v[5] * x2;
    a2 += (TA)v[6] * x0 + (TA)v[7] * x1 + (TA)v[8] * x2;
  }
  a[0] = a0;
  a[1] = a1;
  a[2] = a2;
}

// (K x)[3 row + k] accumulated in TA with the FP64 lane layout, lane = 3 rl + k (row rl of the wavefront's 21,
// output component k); valid on lanes with `out` set. For a fixed j the wavefront's 21 rows read 21 adjacent
// 72-byte blocks (1,512 contiguous bytes), each lane the 24 bytes of its block row. The lane keeps one sum per
// column l, a_l = sum_j K_j[k][l] x_j[l] in block order, and returns a_0 + (a_1 + a_2): the order of the former
// layout with one lane per block entry (7 rows per wavefront, combined by two lane shifts), so y is bitwise the
// same as then and as a CSR SpMV's. The loads of SPMV_UNROLL blocks are issued before their products are summed.
template <typename TV, typename TX, typename TA>
__device__ inline TA spmvRow3(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ sellStart,
                              const int32_t* __restrict__ scol, const TV* __restrict__ val,
                              const TX* __restrict__ x, int64_t nrows, unsigned rb, int64_t& row, int& k,
                              bool& out) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rl = lane / 3;
  k            = lane - 3 * rl;
  row          = ((int64_t)rb * (SPMV_BS / 64) + wave) * spmvRowsPerWave<TV>() + rl;
  out          = rl < spmvRowsPerWave<TV>() && row < nrows;
  if (!out) return TA(0);
  const int64_t base = (int64_t)sellStart[row >> 6] + (row & 63);
  const TV* V        = val + base * 9 + 3 * k;
  const int32_t* C   = scol + base;
  const int n        = rowptr[row + 1] - rowptr[row];
  TA a0 = TA(0), a1 = TA(0), a2 = TA(0);
  int j = 0;
  for (; j + SPMV_UNROLL <= n; j += SPMV_UNROLL) {
    int32_t c[SPMV_UNROLL];
    TV v[SPMV_UNROLL][3];
#pragma unroll
    for (int u = 0; u < SPMV_UNROLL; ++u) {
      c[u] = C[(j + u) * kSellRows];
#pragma unroll
      for (int l = 0; l < 3; ++l) v[u][l] = V[(int64_t)(j + u) * (kSellRows * 9) + l];
    }
#pragma unroll
    for (int u = 0; u < SPMV_UNROLL; ++u) {
      const TX* xc = x + (int64_t)c[u] * 3;
      a0 += (TA)v[u][0] * (TA)xc[0];
      a1 += (TA)v[u][1] * (TA)xc[1];
      a2 += (TA)v[u][2] * (TA)xc[2];
    }
  }
  for (; j < n; ++j) {
    const TV* vj = V + (int64_t)j * (kSellRows * 9);
    const TX* xc = x + (int64_t)C[j * kSellRows] * 3;
    a0 += (TA)vj[0] * (TA)xc[0];
    a1 += (TA)vj[1] * (TA)xc[1];
    a2 += (TA)vj[2] * (TA)xc[2];
  }
  return a0 + (a1 + a2);
}

// rev: workgroup i takes row block gridDim.x - 1 - i (the partial stays at the row block's index, so the fold
// order and y are the same either way). The PCG loops alternate the direction from one iteration to the next:
// the matrix (FP32 259 MB, FP64 518 MB, plus 29 MB of column indices) is larger than the 256 MB last-level
// cache, and a pass in one direction evicts what the next pass in the same direction reads first, while a pass
// in the other direction starts on the part just read (64^3: FP64 134.4 -> 119.0 us, FP32 57.0 -> 53.3 us).
template <typename TV, typename TX, bool DOT>
__global__ __launch_bounds__(SPMV_BS) void spmvKernel(const int32_t* __restrict__ rowptr,
                                                      const int32_t* __restrict__ sellStart,
                                                      const int32_t* __restrict__ scol,
                                                      const TV* __restrict__ val, const TX* __restrict__ x,
                                                      TX* __restrict__ y, int64_t nrows,
                                                      double* __restrict__ partial,
                                                      const PcgScalars* __restrict__ sc, int rev) {
  if (sc && sc->done) return;
  const unsigned rb = rev ? gridDim.x - 1u - blockIdx.x : blockIdx.x;
  double d = 0.0;
  if constexpr (sizeof(TV) == 4) {
    const int64_t row = (int64_t)rb * SPMV_BS + threadIdx.x;
    if (row < nrows) {
      TX a[3];
      rowSell3<TV, TX, TX>(sellStart, scol, val, x, row, a);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        y[row * 3 + k] = a[k];
        if (DOT) d += (double)x[row * 3 + k] * (double)a[k];
      }
    }
  } else {
    int64_t row;
    int k;
    bool out;
    const TX yk = spmvRow3<TV, TX, TX>(rowptr, sellStart, scol, val, x, nrows, rb, row, k, out);
    if (out) {
      y[row * 3 + k] = yk;
      if (DOT) d = (double)x[row * 3 + k] * (double)yk;
    }
  }
  if (DOT) {
    const double s = blockSum<SPMV_BS>(d);
    if (threadIdx.x == 0) partial[rb] = s;
  }
}

// r = b - K x with FP64 accumulation (K may be FP32); used by iterative refinement. Same lane layouts.
template <typename TV>
__global__ __launch_bounds__(SPMV_BS) void residualKernel(const int32_t* __restrict__ rowptr,
                                                          const int32_t* __restrict__ sellStart,
                                                          const int32_t* __restrict__ scol,
                                                          const TV* __restrict__ val,
                                                          const double* __restrict__ x,
                                                          const double* __restrict__ b, double* __restrict__ r,
                                                          int64_t nrows, double* __restrict__ partial) {
  double d = 0.0;
  if constexpr (sizeof(TV) == 4) {
    const int64_t row = (int64_t)blockIdx.x * SPMV_BS + threadIdx.x;
    if (row < nrows) {
      double a[3];
      rowSell3<TV, double, double>(sellStart, scol, val, x, row, a);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double rk = b[row * 3 + k] - a[k];
        r[row * 3 + k]  = rk;
        d += rk * rk;
      }
    }
  } else {
    int64_t row;
    int k;
    bool out;
    const double kx = spmvRow3<TV, double, double>(rowptr, sellStart, scol, val, x, nrows, blockIdx.x, row, k, out);
    if (out) {
      const double rk = b[row * 3 + k] - kx;
      r[row * 3 + k]  = rk;
      d               = rk * rk;
    }
  }
  const double s = blockSum<SPMV_BS>(d);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// partial dot products (grid-stride, fixed grid) — deterministic given the launch shape
__global__ __launch_bounds__(256) void dotKernel(const double* __restrict__ a, const double* __restrict__ b,
                                                 int64_t n, double* __restrict__ partial) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) s += a[i] * b[i];
  s = blockSum<256>(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(1024) void foldKernel(const double* __restrict__ partial, int n, double* __restrict__ out) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) s += partial[i];
  s = blockSum<1024>(s);
  if (threadIdx.x == 0) *out = s;
}

__global__ void axpyKernel(double a, const double* __restrict__ x, double* __restrict__ y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) y[i] += a * x[i];
}

// ---------------- PCG (Jacobi, FP64) ----------------
__global__ void jacobiDiagKernel(const int32_t* __restrict__ diag, const double* __restrict__ K, int64_t nrows,
                                 double* __restrict__ dinv) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nrows * 3) return;
  const int64_t node = i / 3;
  const int k        = (int)(i - node * 3);
  const double d     = K[(int64_t)diag[node] * 9 + k * 4];
  dinv[i]            = d != 0.0 ? 1.0 / d : 1.0;
}

// r = b, x = 0, p = Dinv r; partials of rr and rz
__global__ __launch_bounds__(256) void pcgInitKernel(const double* __restrict__ b, const double* __restrict__ dinv,
                                                     double* __restrict__ x, double* __restrict__ r,
                                                     double* __restrict__ p, int64_t n,
                                                     double* __restrict__ partial /* 2*grid */) {
  double rr = 0.0, rz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double bi = b[i];
    x[i]            = 0.0;
    r[i]            = bi;
    const double z  = dinv[i] * bi;
    p[i]            = z;
    rr += bi * bi;
    rz += bi * z;
  }
  rr = blockSum<256>(rr);
  rz = blockSum<256>(rz);
  if (threadIdx.x == 0) {
    partial[blockIdx.x]             = rr;
    partial[gridDim.x + blockIdx.x] = rz;
  }
}

__global__ __launch_bounds__(1024) void pcgInitFold(const double* __restrict__ partial, int nb, double tol2,
                                                    PcgScalars* __restrict__ sc) {
  double rr = 0.0, rz = 0.0;
  for (int i = threadIdx.x; i < nb; i += 1024) {
    rr += partial[i];
    rz += partial[nb + i];
  }
  rr = blockSum<1024>(rr);
  rz = blockSum<1024>(rz);
  if (threadIdx.x == 0) {
    sc->bb        = rr;
    sc->rr        = rr;
    sc->rz        = rz;
    sc->rz_new    = rz;
    const double thr = tol2 * rr;
    sc->threshold = thr > 2.2250738585072014e-308 ? thr : 2.2250738585072014e-308;
    sc->iters     = 0;
    sc->done      = (rr == 0.0 || rr < sc->threshold) ? 1 : 0;
    sc->alpha = sc->beta = 0.0;
  }
}

// First stage of the p.q fold: block b sums its slice of the SpMV's per-block partials, in order, into
// pq2[b] (kPqBlocks values). A single block summing one partial per SpMV block took 17 us per iteration at
// 1 M elements (38,631 partials); the update kernel folds the kPqBlocks values itself. Launches of at most
// kPqDirect SpMV blocks skip this stage: every update block then folds the SpMV's partials directly.
constexpr int kPqBlocks = 64;
constexpr int kPqDirect = 2048;
__global__ __launch_bounds__(256) void pcgPqFold(const double* __restrict__ partial, int nb,
                                                 double* __restrict__ pq2, const PcgScalars* __restrict__ sc) {
  if (sc->done) return;
  const int chunk = (nb + kPqBlocks - 1) / kPqBlocks;
  const int lo = blockIdx.x * chunk, hi = lo + chunk < nb ? lo + chunk : nb;
  double s = 0.0;
  for (int i = lo + threadIdx.x; i < hi; i += 256) s += partial[i];
  s = blockSum<256>(s);
  if (threadIdx.x == 0) pq2[blockIdx.x] = s;
}

// alpha = r.z / p.q (every block folds the npq values pq[] -- pcgPqFold's kPqBlocks sums or the SpMV's
// partials -- in the same order, so all blocks agree bitwise; block 0 records the scalars), then
// x += alpha p, r -= alpha q with partials of r.r and r.z (z = M^-1 r)
template <typename T>
__global__ __launch_bounds__(256) void pcgUpdateKernel(T* __restrict__ x, T* __restrict__ r,
                                                       const T* __restrict__ p, const T* __restrict__ q,
                                                       const T* __restrict__ pre, int blockPre, int64_t n,
                                                       PcgScalars* __restrict__ sc, const double* __restrict__ pq,
                                                       int npq, double* __restrict__ partial /* 2*grid */) {
  if (sc->done) return;
  __shared__ double alphaS;
  {
    double v = 0.0;
    for (int i = threadIdx.x; i < npq; i += 256) v += pq[i];
    v = blockSum<256>(v);
    if (threadIdx.x == 0) {
      const double rz = sc->rz_new;  // r.z of the current iterate (set by the previous p update)
      alphaS          = rz / v;
      if (blockIdx.x == 0) {
        sc->rz    = rz;
        sc->pq    = v;
        sc->alpha = alphaS;
      }
    }
    __syncthreads();
  }
  const T alpha = (T)alphaS;
  double rr = 0.0, rz = 0.0;
  if (!blockPre) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
      x[i] += alpha * p[i];
      const T ri = r[i] - alpha * q[i];
      r[i]       = ri;
      rr += (double)ri * (double)ri;
      rz += (double)ri * (double)(pre[i] * ri);
    }
  } else {
    const int64_t nn = n / 3;
    for (int64_t nd = (int64_t)blockIdx.x * 256 + threadIdx.x; nd < nn; nd += (int64_t)gridDim.x * 256) {
      T rl[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int64_t i = nd * 3 + k;
        x[i] += alpha * p[i];
        rl[k] = r[i] - alpha * q[i];
        r[i]  = rl[k];
        rr += (double)rl[k] * (double)rl[k];
      }
      const T* B = pre + nd * 9;
#pragma unroll
      for (int k = 0; k < 3; ++k) rz += (double)rl[k] * (double)(B[k * 3] * rl[0] + B[k * 3 + 1] * rl[1] + B[k * 3 + 2] * rl[2]);
    }
  }
  rr = blockSum<256>(rr);
  rz = blockSum<256>(rz);
  if (threadIdx.x == 0) {
    partial[blockIdx.x]             = rr;
    partial[gridDim.x + blockIdx.x] = rz;
  }
}

// Convergence test, beta and p = M^-1 r + beta p in one launch. Every block folds the update kernel's
// partials (rr, rz) itself in the same fixed order, so all blocks agree bitwise; block 0 records the
// scalars. beta = rz_new / rz uses the r.z of the previous iterate, which only pcgAlphaFold writes.
template <typename T>
__global__ __launch_bounds__(256) void pcgBetaPKernel(T* __restrict__ p, const T* __restrict__ r,
                                                      const T* __restrict__ pre, int blockPre, int64_t n,
                                                      const double* __restrict__ partial, int nb,
                                                      PcgScalars* __restrict__ sc, int maxIter) {
  if (sc->done) return;
  __shared__ double bc[2];
  double rr = 0.0, rz = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) {
    rr += partial[i];
    rz += partial[nb + i];
  }
  rr = blockSum<256>(rr);
  rz = blockSum<256>(rz);
  if (threadIdx.x == 0) {
    bc[0] = rr;
    bc[1] = rz;
  }
  __syncthreads();
  rr              = bc[0];
  rz              = bc[1];
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
  const T beta = (T)bt;
  if (!blockPre) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
      p[i] = pre[i] * r[i] + beta * p[i];
  } else {
    const int64_t nn = n / 3;
    for (int64_t nd = (int64_t)blockIdx.x * 256 + threadIdx.x; nd < nn; nd += (int64_t)gridDim.x * 256) {
      const T r0 = r[nd * 3], r1 = r[nd * 3 + 1], r2 = r[nd * 3 + 2];
      const T* B = pre + nd * 9;
#pragma unroll
      for (int k = 0; k < 3; ++k) p[nd * 3 + k] = B[k * 3] * r0 + B[k * 3 + 1] * r1 + B[k * 3 + 2] * r2 + beta * p[nd * 3 + k];
    }
  }
}

// ---------------- FP32 block-Jacobi helpers ----------------
__global__ void blockInvKernel(const int32_t* __restrict__ diag, const float* __restrict__ K, int64_t nrows,
                               float* __restrict__ binv) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nrows) return;
  const float* A = K + (int64_t)diag[i] * 9;
  // invert in double for robustness, store FP32
  const double a00 = A[0], a01 = A[1], a02 = A[2], a10 = A[3], a11 = A[4], a12 = A[5], a20 = A[6], a21 = A[7],
               a22 = A[8];
  const double det = a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20) + a02 * (a10 * a21 - a11 * a20);
  float* B = binv + i * 9;
  if (det == 0.0) {
    for (int k = 0; k < 9; ++k) B[k] = (k % 4 == 0) ? 1.0f : 0.0f;
    return;
  }
  const double id = 1.0 / det;
  B[0] = (float)((a11 * a22 - a12 * a21) * id);
  B[1] = (float)((a02 * a21 - a01 * a22) * id);
  B[2] = (float)((a01 * a12 - a02 * a11) * id);
  B[3] = (float)((a12 * a20 - a10 * a22) * id);
  B[4] = (float)((a00 * a22 - a02 * a20) * id);
  B[5] = (float)((a02 * a10 - a00 * a12) * id);
  B[6] = (float)((a10 * a21 - a11 * a20) * id);
  B[7] = (float)((a01 * a20 - a00 * a21) * id);
  B[8] = (float)((a00 * a11 - a01 * a10) * id);
}

// FP32 PCG init from an FP64 right-hand side scaled by 1/s (s = ||r||, keeps FP32 in range)
__global__ __launch_bounds__(256) void pcgInitF32(const double* __restrict__ b, double invScale,
                                                  const float* __restrict__ binv, float* __restrict__ x,
                                                  float* __restrict__ r, float* __restrict__ p, int64_t nn,
                                                  double* __restrict__ partial) {
  double rr = 0.0, rz = 0.0;
  for (int64_t nd = (int64_t)blockIdx.x * 256 + threadIdx.x; nd < nn; nd += (int64_t)gridDim.x * 256) {
    float rl[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      rl[k]         = (float)(b[nd * 3 + k] * invScale);
      r[nd * 3 + k] = rl[k];
      x[nd * 3 + k] = 0.0f;
      rr += (double)rl[k] * (double)rl[k];
    }
    const float* B = binv + nd * 9;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float z = B[k * 3] * rl[0] + B[k * 3 + 1] * rl[1] + B[k * 3 + 2] * rl[2];
      p[nd * 3 + k] = z;
      rz += (double)rl[k] * (double)z;
    }
  }
  rr = blockSum<256>(rr);
  rz = blockSum<256>(rz);
  if (threadIdx.x == 0) {
    partial[blockIdx.x]             = rr;
    partial[gridDim.x + blockIdx.x] = rz;
  }
}

// two-level form: r = b / s, x = 0 and r.r partials (z and p follow from the coarse correction); clears the done
// flag, which the coarse kernels read before the init fold sets it
__global__ __launch_bounds__(256) void pcgInitR32(const double* __restrict__ b, double invScale,
                                                  float* __restrict__ x, float* __restrict__ r, int64_t n,
                                                  double* __restrict__ partial, PcgScalars* __restrict__ sc) {
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->done = 0;
  double rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const float ri = (float)(b[i] * invScale);
    r[i]           = ri;
    x[i]           = 0.0f;
    rr += (double)ri * (double)ri;
  }
  rr = blockSum<256>(rr);
  if (threadIdx.x == 0) partial[blockIdx.x] = rr;
}

// x64 += scale * x32
__global__ void accumulateF32(double* __restrict__ x, const float* __restrict__ d, double scale, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    x[i] += scale * (double)d[i];
}

inline int vecGrid(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > kReduceBlocks) g = kReduceBlocks;
  return (int)(g < 1 ? 1 : g);
}

}  // namespace

static double* pq2(ik_handle* h) { return h->partial.p + h->partial.n - 8 - kPqBlocks; }
// the PCG SpMV's per-block p.q partials, apart from the vector kernels' partials at the start (the update
// kernel reads the former while its blocks write the latter)
static double* pqPartials(ik_handle* h) { return h->partial.p + 2 * kReduceBlocks; }

// partial[] holds one launch's per-block partials at its start, the PCG SpMV's partials from 2 kReduceBlocks
// on, the p.q first-stage sums (kPqBlocks) in a region of their own before the last 8 entries, and fold
// results in its last entry
static void ensureWork(ik_handle* h) {
  const int64_t n = h->ndof;
  const int64_t spmvBlocks = spmvGrid<double>(h->nn);  // the FP64 layout launches the most blocks
  h->partial.ensure((size_t)(2 * kReduceBlocks + spmvBlocks) + kPqBlocks + 8);
  if (!h->sc.p) {
    h->sc.ensure(1);
    IK_HIP(hipHostMalloc((void**)&h->hsc, sizeof(PcgScalars), hipHostMallocDefault));
    IK_HIP(hipHostMalloc((void**)&h->hred, sizeof(double) * 4, hipHostMallocDefault));
  }
  (void)n;
}

void spmv64(ik_handle* h, const double* x, double* y) {
  ensureWork(h);
  const unsigned grid = spmvGrid<double>(h->nn);
  ScopedKernelTimer t(h, KSPMV);
  hipLaunchKernelGGL((spmvKernel<double, double, false>), dim3(grid), dim3(SPMV_BS), 0, h->stream, h->rowptr.p,
                     h->sellStart.p, h->scol.p, h->K64.p, x, y, h->nn, nullptr, nullptr, 0);
}

void spmv32(ik_handle* h, const float* x, float* y) {
  ensureWork(h);
  const unsigned grid = spmvGrid<float>(h->nn);
  ScopedKernelTimer t(h, KSPMV);
  hipLaunchKernelGGL((spmvKernel<float, float, false>), dim3(grid), dim3(SPMV_BS), 0, h->stream, h->rowptr.p,
                     h->sellStart.p, h->scol.p, h->K32.p, x, y, h->nn, nullptr, nullptr, 0);
}

double dot(ik_handle* h, const double* a, const double* b, int64_t n) {
  ensureWork(h);
  const int g = vecGrid(n);
  hipLaunchKernelGGL(dotKernel, dim3(g), dim3(256), 0, h->stream, a, b, n, h->partial.p);
  hipLaunchKernelGGL(foldKernel, dim3(1), dim3(1024), 0, h->stream, h->partial.p, g, h->partial.p + h->partial.n - 1);
  IK_HIP(hipMemcpyAsync(h->hred, h->partial.p + h->partial.n - 1, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  IK_HIP(hipStreamSynchronize(h->stream));
  finishErrCheck(h);
  return h->hred[0];
}

double* partialsFor(ik_handle* h, int nb) {
  ensureWork(h);
  require(nb >= 1 && nb <= kReduceBlocks, IK_EINVAL, "internal error: too many reduction partials");
  return h->partial.p;
}

double foldPartials(ik_handle* h, int nb) {
  hipLaunchKernelGGL(foldKernel, dim3(1), dim3(1024), 0, h->stream, h->partial.p, nb, h->partial.p + h->partial.n - 1);
  IK_HIP(hipMemcpyAsync(h->hred, h->partial.p + h->partial.n - 1, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  IK_HIP(hipStreamSynchronize(h->stream));
  finishErrCheck(h);
  return h->hred[0];
}

void axpy(ik_handle* h, double a, const double* x, double* y, int64_t n) {
  hipLaunchKernelGGL(axpyKernel, dim3(vecGrid(n)), dim3(256), 0, h->stream, a, x, y, n);
}

void residualBlock64(ik_handle* h, const double* b, const double* x, double* r) {
  ensureWork(h);
  const unsigned grid = spmvGrid<double>(h->nn);
  hipLaunchKernelGGL((residualKernel<double>), dim3(grid), dim3(SPMV_BS), 0, h->stream, h->rowptr.p, h->sellStart.p,
                     h->scol.p, h->K64.p, x, b, r, h->nn, h->partial.p);
}

// Poll interval: iterations enqueued before the first host check of the done flag (32: a two-level inner solve
// takes 24-32 at 64^3, so most need no second check) and between later checks (64). Iterations enqueued after
// convergence return at once (a few us of launches each); a check leaves the device idle for a host round trip
// (20-45 us) while the next iterations are enqueued.
static int pollChunk(int enq) { return enq == 0 ? 32 : 64; }

void pcgJacobi64(ik_handle* h, const double* b, double* x, double rtol, int maxIter, int* iters, double* relres) {
  ensureWork(h);
  const int64_t n = h->ndof;
  h->pr.ensure(n);
  h->pp.ensure(n);
  h->pq.ensure(n);
  h->pdiag.ensure(n);
  const int g = vecGrid(n);
  const unsigned sgrid = spmvGrid<double>(h->nn);
  hipLaunchKernelGGL(jacobiDiagKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->diagBlk.p,
                     h->K64.p, h->nn, h->pdiag.p);
  hipLaunchKernelGGL(pcgInitKernel, dim3(g), dim3(256), 0, h->stream, b, h->pdiag.p, x, h->pr.p, h->pp.p, n,
                     h->partial.p);
  hipLaunchKernelGGL(pcgInitFold, dim3(1), dim3(1024), 0, h->stream, h->partial.p, g, rtol * rtol, h->sc.p);
  if (maxIter <= 0) maxIter = (int)(2 * n);
  int enq = 0;
  for (;;) {
    for (int c = 0, cn = pollChunk(enq); c < cn && enq < maxIter; ++c, ++enq) {
      {
        ScopedKernelTimer t(h, KSPMV);
        hipLaunchKernelGGL((spmvKernel<double, double, true>), dim3(sgrid), dim3(SPMV_BS), 0, h->stream, h->rowptr.p,
                           h->sellStart.p, h->scol.p, h->K64.p, h->pp.p, h->pq.p, h->nn, pqPartials(h), h->sc.p,
                           enq & 1);
      }
      const double* pqSrc = pqPartials(h);
      int npq             = (int)sgrid;
      if (sgrid > (unsigned)kPqDirect) {
        hipLaunchKernelGGL(pcgPqFold, dim3(kPqBlocks), dim3(256), 0, h->stream, pqPartials(h), (int)sgrid, pq2(h),
                           h->sc.p);
        pqSrc = pq2(h);
        npq   = kPqBlocks;
      }
      hipLaunchKernelGGL((pcgUpdateKernel<double>), dim3(g), dim3(256), 0, h->stream, x, h->pr.p, h->pp.p, h->pq.p,
                         h->pdiag.p, 0, n, h->sc.p, pqSrc, npq, h->partial.p);
      hipLaunchKernelGGL((pcgBetaPKernel<double>), dim3(g), dim3(256), 0, h->stream, h->pp.p, h->pr.p, h->pdiag.p, 0,
                         n, h->partial.p, g, h->sc.p, maxIter);
    }
    IK_HIP(hipMemcpyAsync(h->hsc, h->sc.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, h->stream));
    IK_HIP(hipStreamSynchronize(h->stream));
    finishErrCheck(h);
    if (h->hsc->done || enq >= maxIter) break;
  }
  *iters  = h->hsc->iters;
  *relres = h->hsc->bb > 0 ? std::sqrt(h->hsc->rr / h->hsc->bb) : 0.0;
}

void pcgMixed(ik_handle* h, const double* b, double* x, const ik_solver_opts* o, ik_solver_info* info,
              double bnormKnown) {
  ensureWork(h);
  const bool twoLevel = o->kind == IK_PCG_TWOLEVEL_FP32_IR;
  const int64_t n = h->ndof;
  h->fr.ensure(n);
  h->fp.ensure(n);
  h->fq.ensure(n);
  h->fx.ensure(n);
  h->fdiag.ensure((size_t)h->nn * 9);
  h->pr.ensure(n);
  const int g = vecGrid(n);
  // vector kernels over nodes: 512 blocks (the p update folds 2 partials per block in every block; 512 vs
  // 1024 blocks measured 2.28 vs 2.31 s on the 64^3 mixed Newton path)
  const int gn = std::min(vecGrid(h->nn), 512);
  const unsigned sgrid = spmvGrid<float>(h->nn);
  hipLaunchKernelGGL(blockInvKernel, dim3((unsigned)((h->nn + 255) / 256)), dim3(256), 0, h->stream, h->diagBlk.p,
                     h->K32.p, h->nn, h->fdiag.p);
  if (twoLevel) coarsePrepare(h);
  // x = 0, r = b
  launchZero(x, n, h->stream);
  IK_HIP(hipMemcpyAsync(h->pr.p, b, n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  const double bnorm = bnormKnown >= 0.0 ? bnormKnown : std::sqrt(dot(h, b, b, n));
  info->iterations  = 0;
  info->refinements = 0;
  info->converged   = 0;
  if (bnorm == 0.0) {
    info->rel_residual = 0.0;
    info->converged    = 1;
    return;
  }
  const int maxIter = o->max_iter > 0 ? o->max_iter : (int)(2 * n);
  double rnorm      = bnorm;
  const int maxRef  = o->max_refine > 0 ? o->max_refine : 10;
  for (int sweep = 0; sweep < maxRef; ++sweep) {
    // inner FP32 PCG on K32 d = r / ||r||, to half the reduction the outer tolerance still needs (an inexact
    // Newton step asks for little), but never below inner_rtol
    const double innerTol = std::max(o->inner_rtol, 0.5 * o->rtol * bnorm / rnorm);
    if (twoLevel) {
      hipLaunchKernelGGL(pcgInitR32, dim3(gn), dim3(256), 0, h->stream, h->pr.p, 1.0 / rnorm, h->fx.p, h->fr.p, n,
                         h->partial.p, h->sc.p);
      coarseApply(h, h->fr.p, h->fp.p, nullptr, h->fdiag.p, gn, h->partial.p);  // p = z
    } else {
      hipLaunchKernelGGL(pcgInitF32, dim3(gn), dim3(256), 0, h->stream, h->pr.p, 1.0 / rnorm, h->fdiag.p, h->fx.p,
                         h->fr.p, h->fp.p, h->nn, h->partial.p);
    }
    hipLaunchKernelGGL(pcgInitFold, dim3(1), dim3(1024), 0, h->stream, h->partial.p, gn, innerTol * innerTol,
                       h->sc.p);
    int enq = 0;
    for (;;) {
      for (int c = 0, cn = pollChunk(enq); c < cn && enq < maxIter; ++c, ++enq) {
        {
          ScopedKernelTimer t(h, KSPMV);
          hipLaunchKernelGGL((spmvKernel<float, float, true>), dim3(sgrid), dim3(SPMV_BS), 0, h->stream,
                             h->rowptr.p, h->sellStart.p, h->scol.p, h->K32.p, h->fp.p, h->fq.p, h->nn,
                             pqPartials(h), h->sc.p, enq & 1);
        }
        const double* pqSrc = pqPartials(h);
        int npq             = (int)sgrid;
        if (sgrid > (unsigned)kPqDirect) {
          hipLaunchKernelGGL(pcgPqFold, dim3(kPqBlocks), dim3(256), 0, h->stream, pqPartials(h), (int)sgrid,
                             pq2(h), h->sc.p);
          pqSrc = pq2(h);
          npq   = kPqBlocks;
        }
        if (twoLevel) {
          coarseIteration(h, pqSrc, npq, h->fx.p, h->fr.p, h->fp.p, h->fq.p, h->fdiag.p, gn, h->partial.p, maxIter);
          continue;
        }
        hipLaunchKernelGGL((pcgUpdateKernel<float>), dim3(gn), dim3(256), 0, h->stream, h->fx.p, h->fr.p, h->fp.p,
                           h->fq.p, h->fdiag.p, 1, n, h->sc.p, pqSrc, npq, h->partial.p);
        hipLaunchKernelGGL((pcgBetaPKernel<float>), dim3(gn), dim3(256), 0, h->stream, h->fp.p, h->fr.p, h->fdiag.p,
                           1, n, h->partial.p, gn, h->sc.p, maxIter);
      }
      IK_HIP(hipMemcpyAsync(h->hsc, h->sc.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, h->stream));
      IK_HIP(hipStreamSynchronize(h->stream));
      finishErrCheck(h);
      if (h->hsc->done || enq >= maxIter) break;
    }
    info->iterations += h->hsc->iters + 1;
    info->refinements += 1;
    // x += ||r|| d ; r = b - K x in FP64
    hipLaunchKernelGGL(accumulateF32, dim3(g), dim3(256), 0, h->stream, x, h->fx.p, rnorm, n);
    hipLaunchKernelGGL((residualKernel<float>), dim3(sgrid), dim3(SPMV_BS), 0, h->stream, h->rowptr.p, h->sellStart.p,
                       h->scol.p, h->K32.p, x, b, h->pr.p, h->nn, h->partial.p);
    hipLaunchKernelGGL(foldKernel, dim3(1), dim3(1024), 0, h->stream, h->partial.p, (int)sgrid,
                       h->partial.p + h->partial.n - 1);
    IK_HIP(hipMemcpyAsync(h->hred, h->partial.p + h->partial.n - 1, sizeof(double), hipMemcpyDeviceToHost,
                          h->stream));
    IK_HIP(hipStreamSynchronize(h->stream));
    finishErrCheck(h);
    rnorm              = std::sqrt(h->hred[0]);
    info->rel_residual = rnorm / bnorm;
    if (info->rel_residual <= o->rtol) {
      info->converged = 1;
      break;
    }
  }
  if (twoLevel) coarseRecord(h, info->iterations, info->rel_residual);
}

}  // namespace ik
