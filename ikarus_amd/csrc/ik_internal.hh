// ik_internal.hh — handle state and shared helpers of the MI355X hot-path library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/ikarus_amd.h"

namespace ik {

struct Mlp;
struct Coarse;

// Allocator whose value-less construct default-initialises: a HostVec made with a size (or resized) leaves its
// elements unwritten, so the pattern build's large arrays are first touched by the parallel loops that fill
// them instead of being zeroed on one thread beforehand. Construction with a value still writes it.
template <class T>
struct NoInitAlloc : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() noexcept = default;
  template <class U>
  NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
  template <class U>
  void construct(U* p) noexcept(std::is_nothrow_default_constructible<U>::value) {
    ::new ((void*)p) U;
  }
  template <class U, class... A>
  void construct(U* p, A&&... a) {
    ::new ((void*)p) U(std::forward<A>(a)...);
  }
};
template <class T>
using HostVec = std::vector<T, NoInitAlloc<T>>;

struct Error {
  int code;
  std::string msg;
};

#define IK_HIP(x)                                                                                  \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess) throw ::ik::Error{IK_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)}; \
  } while (0)

inline void require(bool ok, int code, const std::string& msg) {
  if (!ok) throw Error{code, msg};
}
// literal messages: the string is built only on failure (the pattern checks call this millions of times)
inline void require(bool ok, int code, const char* msg) {
  if (!ok) throw Error{code, msg};
}

// Owning device buffer. Grows only; contents are not preserved on growth.
template <class T>
struct DevBuf {
  T* p     = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void ensure(size_t count) {
    if (count <= n && p) return;
    release();
    IK_HIP(hipMalloc(&p, (count ? count : 1) * sizeof(T)));
    n = count;
  }
  void upload(const T* src, size_t count, hipStream_t s) {
    ensure(count);
    if (count) IK_HIP(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
};

// Device scalars of one PCG run (Eigen ConjugateGradient state, linearsolver.cpp:23-24)
struct PcgScalars {
  double rz;        // r . z of the current iterate
  double rz_new;
  double pq;        // p . (K p)
  double rr;        // r . r
  double alpha, beta;
  double threshold; // tol^2 ||b||^2
  int32_t done;
  int32_t iters;
  double bb;        // ||b||^2
};

enum KernelId { KASM_KR = 0, KASM_R = 1, KSPMV = 2, KMLP = 3, KGATHER = 4, KCOUNT = 5 };

// Reduction partials: one double per block, final fold in fixed order (deterministic).
constexpr int kReduceBlocks = 1024;

// elements per tile of the plain element's tangent kernel (one workgroup, 4 elements per 64-lane wavefront)
#ifndef IK_TILE_ELEMS
#define IK_TILE_ELEMS 16
#endif
// block rows per slice of the sliced-ELL storage of K
constexpr int kSellRows = 64;

constexpr int kTileElems   = IK_TILE_ELEMS;
constexpr int kTileThreads = kTileElems * 16;
constexpr int kTileEntries = kTileElems * 36;  // upper element blocks per tile
static_assert(kTileElems % 4 == 0 && kTileElems <= 16, "tiles hold whole wavefronts, entries fit 10 bits");

}  // namespace ik

struct ik_handle {
  std::string err;
  int64_t bad_elem = -1;
  hipStream_t stream = nullptr;   // all work is ordered on this stream
  hipStream_t stream2 = nullptr;  // assembly gathers, joined back into `stream` by events
  std::vector<hipEvent_t> chunkEvents;
  hipEvent_t joinEvent = nullptr;

  // problem
  // Plane strain (dim 2): the Q1 mesh runs as its one-layer hex8 extrusion of thickness sqrt(2) with u_z = 0 on
  // every node; node v of the 2D mesh is 3D node v (z = 0) and v + nn2 (z = sqrt(2)). A displacement independent
  // of z has E33 = E13 = E23 = 0 and the 2-point Gauss rule in z is exact for it, so each layer carries
  // sqrt(2)/2 = 1/sqrt(2) of the 2D residual and the 3D residual norm equals the 2D one.
  bool plane2d = false;
  int64_t nn2 = 0;                  // 2D nodes
  std::vector<uint8_t> hfixed2;     // 2D Dirichlet flags [nn2 * 2]
  std::vector<int32_t> hconn2;      // 2D connectivity [ne * 4]
  ik::DevBuf<double> map2d;         // 2D-sized staging for device-pointer arguments
  int64_t nn = 0, ne = 0, ndof = 0;
  int material = IK_NEOHOOKE;
  double lam = 0, mu = 0;
  int neas = 0;
  std::vector<double> hX;
  std::vector<int32_t> hconn;
  std::vector<uint8_t> hfixed;
  int64_t nfixed = 0;
  ik::DevBuf<double> X;
  ik::DevBuf<int32_t> conn;
  ik::DevBuf<uint8_t> fixed;
  ik::DevBuf<int32_t> fixedList;  // constrained DOF indices
  // inhomogeneous Dirichlet values u_D(lambda) = lambda * ubar on constrained DOFs (none unless set)
  bool hasUbar = false;
  ik::DevBuf<double> ubar, idbcF;
  // arc length: the step's accumulated increment, the two correction directions and the load direction
  ik::DevBuf<double> arcDD, arcDR, arcDL, arcF, arcRhs;  // arcRhs: Fext0 - F_idbc
  // loads
  int64_t nfaces = 0;
  ik::DevBuf<uint8_t> faceBits;  // per element: bit 2*axis + side set for each loaded face (empty: none)
  double trac[3] = {0, 0, 0};
  int64_t npl = 0;
  ik::DevBuf<int64_t> plDof;
  ik::DevBuf<double> plCoef;
  std::vector<int64_t> hplDof;
  std::vector<double> hplCoef;

  // 3x3 block CSR pattern (block row = node)
  bool patternBuilt = false;
  ik::HostVec<int32_t> hrowptr, hcolidx, hdiag;
  int64_t nnzb = 0;
  ik::DevBuf<int32_t> rowptr, colidx, diagBlk;
  // the upper blocks (column >= row) the gather assembles, in gather order (staging chunk by staging chunk, by
  // storage slot inside a chunk): the storage slot of each and of its transpose; chunkQ[c] = first of chunk c
  std::vector<int64_t> chunkQ;
  ik::DevBuf<int32_t> upBlk, upT;
  // FULL Dirichlet treatment per upper block: bit t (t = 3k+l) zeroes entry t (constrained row or column),
  // bit 9+k sets diagonal entry (k,k) of a diagonal block to one
  ik::DevBuf<uint16_t> upMask;
  // Tangent staging: slots ordered upper block by upper block in gather order (segptr[q] = first slot of the
  // q-th gathered upper block),
  // each segment in a fixed order, so the gather reads contiguous runs and sums them deterministically.
  //  EAS element: one slot per element block; bdest[e*36+sb] = (slot << 1) | transposed, segments ordered
  //  by element.
  //  Plain element (tiled): elements grouped in tiles of kTileElems (Morton order of the element centroids);
  //  the blocks of a tile that fall on the same upper block form a group, summed in the tile. An upper
  //  block touched by one tile only is then complete and written straight into K; the others get one
  //  staging slot per group (segments ordered by tile). tileElem[t*16+s] = element (-1: padding of the
  //  last tile); tgptr[t] = first group of tile t (groups in gather order); gdst[g] = the slot of
  //  group g, or -(K slot + 1) for a direct group, whose mirror slot is gpt[g] and FULL mask gmsk[g].
  //  epos[t*576 + j] = group | rank << 10 | transposed << 15 of the tile's entry j = 36 s + sb (rank = its
  //  position in the group by increasing j; padding entries target spare accumulators past the groups),
  //  tmaxr[t] = the tile's largest group size.
  // Residual staging: node by node, each node's (element, corner) incidences ordered by element;
  // rdest[e*8+a] = slot, ncptr[n] = first slot of node n
  ik::DevBuf<int32_t> bdest, segptr, ncptr, rdest;
  bool tiled = false;
  int64_t ntiles = 0, nstage = 0;  // tiles; staging slots of the tangent
  ik::DevBuf<int32_t> tileElem, tgptr, gdst, gpt;
  ik::DevBuf<uint16_t> epos, gmsk;
  ik::DevBuf<uint8_t> tmaxr;
  // staging chunks (tiles when tiled, elements otherwise) and the block rows complete after each chunk
  std::vector<int64_t> chunkElem, chunkRow;
  // element-by-element staging of K_e (36 upper 3x3 blocks) and R_e (24)
  ik::DevBuf<double> ebeK64, ebeR;
  ik::DevBuf<float> ebeK32;
  // K storage: sliced ELL over the block rows. Rows come in slices of kSellRows; slice s holds (its longest row) x
  // kSellRows slots, slot sellStart[s] + j * kSellRows + r is block j (in column order) of the slice's row r, so for
  // a fixed j the slice's rows are adjacent. Each slot holds one 3x3 block, row-major (9 values). Block b of the
  // CSR pattern (rowptr / colidx order) lives in slot hpos[b]; scol[slot] is its column (padding slots: the row
  // itself, with zero values). Every K-addressing index array (diagBlk, upBlk, upT, the direct groups' gdst and
  // gpt) holds slots.
  int64_t nsell = 0;
  ik::HostVec<int32_t> hpos;
  ik::DevBuf<int32_t> sellStart, scol;
  ik::DevBuf<int32_t> blockPos;  // hpos on the device, uploaded when the block view asks for it
  ik::DevBuf<double> K64;
  ik::DevBuf<float> K32;
  int kPrecision = -1, kDbc = -1;
  bool kValid = false;

  // requirement and vectors
  bool bound = false;
  double lambda = 0;
  ik::DevBuf<double> u, R, du, rhs, tmp;

  // EAS: internal variables and, per element, the alpha-update cache of the last staged state: with
  // D = L_D diag(d) L_D^T, diag(d)^{-1} L_D^{-1} L (21x24, column by column), diag(d)^{-1} L_D^{-1} Rt and L_D
  ik::DevBuf<double> alpha, easCache;
  bool easCacheValid = false;

  // error flag (lowest element index with det C <= 0)
  ik::DevBuf<unsigned long long> errFlag;
  // deferErr: inside a Newton loop an assembly enqueues its error-flag copy without waiting for it, and the next
  // synchronisation of `stream` inspects it (errPending until then)
  bool deferErr = false, errPending = false;

  // solver workspace
  ik::DevBuf<double> pr, pz, pp, pq, px, pdiag, pb, partial;
  ik::DevBuf<float> fr, fp, fq, fx, fdiag;  // FP32 PCG vectors; fdiag holds 3x3 block inverses
  ik::DevBuf<ik::PcgScalars> sc;
  ik::PcgScalars* hsc = nullptr;  // pinned host mirror
  double* hred = nullptr;         // pinned host scalar for reductions
  unsigned long long* herr = nullptr;  // pinned host copy of the determinant-failure flag

  // two-level preconditioner of the FP32 inner PCG (coarse.hip)
  ik::Coarse* coarse = nullptr;

  // learned predictor
  ik::Mlp* mlp = nullptr;
  bool mlpLoaded = false;
  // mesh lattice (set by the pattern build): node bounds and the median element extent per axis; the Morton
  // order of the element tiles and the boxes of the coarse aggregates are cells of it
  double latLo[3] = {0, 0, 0}, latHi[3] = {0, 0, 0}, latStep[3] = {1, 1, 1};
  double mlpTraction[3] = {0, 0, 0};
  double mlpCharLen = 1.0;

  // timing
  bool timing = false;
  int64_t tLaunches[ik::KCOUNT] = {};
  double tMs[ik::KCOUNT] = {};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace ik {
// pattern.cpp
void buildPattern(ik_handle* h);
// assemble.hip
void launchZero(double* p, int64_t n, hipStream_t s);
void launchZeroF(float* p, int64_t n, hipStream_t s);
void assembleKR(ik_handle* h, int dbc, int precision, bool wantK, bool wantR);
// throws the error an assembly's deferred flag copy reported; call after a synchronisation of h->stream that
// followed the copy, or with sync = true
void finishErrCheck(ik_handle* h, bool sync = false);
void elementHost(ik_handle* h, int64_t e, double* Ke, double* Re);
void easUpdate(ik_handle* h, const double* duDev);
// element blocks go to staging slots dest[e*36+sb], corner residuals to rdest[e*8+a] (nullptr: element order)
void easStage(ik_handle* h, int64_t e0, int64_t count, bool wantK, const int32_t* dest, const int32_t* rdest);
// sparse.hip
void spmv64(ik_handle* h, const double* x, double* y);
void spmv32(ik_handle* h, const float* x, float* y);
double dot(ik_handle* h, const double* a, const double* b, int64_t n);
void axpy(ik_handle* h, double a, const double* x, double* y, int64_t n);  // y += a x
// per-block partial sums (at most kReduceBlocks) folded in a fixed order
double* partialsFor(ik_handle* h, int nb);
double foldPartials(ik_handle* h, int nb);
double energy(ik_handle* h);  // potential energy at the bound requirement (plain element)
void stressAt(ik_handle* h, int kind, const double xi[3], double* out);  // out: device, ne x 6
void pcgJacobi64(ik_handle* h, const double* b, double* x, double rtol, int maxIter, int* iters, double* relres);
// bnorm: ||b|| when the caller has it (the Newton residual's norm), negative to compute it here
void pcgMixed(ik_handle* h, const double* b, double* x, const ik_solver_opts* o, ik_solver_info* info,
              double bnorm = -1.0);
// coarse.hip
void coarseDestroy(Coarse* c);
void coarsePrepare(ik_handle* h);
void coarseRecord(ik_handle* h, int iterations, double relres);
int coarseBuilds(const ik_handle* h);
void coarseApply(ik_handle* h, const float* r, float* z, float* p, const float* binv, int gn, double* partial);
void coarseIteration(ik_handle* h, const double* pq, int npq, float* x, float* r, float* p, const float* q,
                     const float* binv, int gn, double* partial, int maxIter);
int64_t coarseDofs(const ik_handle* h);
void residualBlock64(ik_handle* h, const double* b, const double* x, double* r);  // r = b - K x (FP64 K)
// mlp.hip
Mlp* mlpCreate(hipStream_t s);
void mlpDestroy(Mlp* m);
int64_t mlpNumParams();
void mlpSetParams(Mlp* m, const float* p, int where, hipStream_t s);
void mlpGetParams(Mlp* m, float* p, int where, hipStream_t s);
void mlpPredict(ik_handle* h, const double* duN, const double* duN1, double lamNext, double dlam, double dlamPrev,
                double dlamPrev2);

// timing helpers (events on the handle's stream around one launch)
// Traction skill of one element (traction.hh:106-138): for each loaded face (bit 2*axis + side of `bits`) the
// one-point rule at the face centre adds -N_a t |J_face| to R_e[3a+k]. Returns that sum for corner a, component
// k (t = the traction component at the current load factor).
__device__ inline double faceLoad(const double* __restrict__ X, const int32_t* __restrict__ conn, int64_t e,
                                  unsigned bits, int a, double t) {
  double out = 0.0;
  for (int f = 0; f < 6; ++f) {
    const int ax = f >> 1, sd = f & 1;
    if (!((bits >> f) & 1) || ((a >> ax) & 1) != sd) continue;  // N_a vanishes on faces without corner a
    double xi[3];
    for (int k = 0; k < 3; ++k) xi[k] = k == ax ? (double)sd : 0.5;
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int b = 0; b < 8; ++b) {
      double fb[3], dN[3];
      for (int k = 0; k < 3; ++k) fb[k] = ((b >> k) & 1) ? xi[k] : 1.0 - xi[k];
      for (int k = 0; k < 3; ++k) {
        const double s = ((b >> k) & 1) ? 1.0 : -1.0;
        dN[k]          = s * fb[(k + 1) % 3] * fb[(k + 2) % 3];
      }
      const int64_t nd = conn[e * 8 + b];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) J[i][j] += X[nd * 3 + i] * dN[j];
    }
    // the two tangent columns of J (axes other than ax, in increasing order)
    double u[3], w[3];
    for (int i = 0; i < 3; ++i) {
      u[i] = ax == 0 ? J[i][1] : J[i][0];
      w[i] = ax == 2 ? J[i][1] : J[i][2];
    }
    const double c0 = u[1] * w[2] - u[2] * w[1];
    const double c1 = u[2] * w[0] - u[0] * w[2];
    const double c2 = u[0] * w[1] - u[1] * w[0];
    double N        = 1.0;
    for (int k = 0; k < 3; ++k) N *= ((a >> k) & 1) ? xi[k] : 1.0 - xi[k];
    out -= N * t * sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  }
  return out;
}

// traction vector at the bound load factor
inline double3 loadedTraction(const ik_handle* h) {
  return make_double3(h->trac[0] * h->lambda, h->trac[1] * h->lambda, h->trac[2] * h->lambda);
}

// HIP events around `launches` back-to-back launches of one kernel on stream s (timing mode only)
struct ScopedKernelTimer {
  ik_handle* h;
  int id;
  hipStream_t s;
  int64_t launches;
  ScopedKernelTimer(ik_handle* hh, int kid, hipStream_t ss = nullptr, int64_t n = 1)
      : h(hh), id(kid), s(ss ? ss : hh->stream), launches(n) {
    if (h->timing) IK_HIP(hipEventRecord(h->ev0, s));
  }
  ~ScopedKernelTimer() {
    if (!h->timing) return;
    (void)hipEventRecord(h->ev1, s);
    (void)hipEventSynchronize(h->ev1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, h->ev0, h->ev1);
    h->tMs[id] += ms;
    h->tLaunches[id] += launches;
  }
};
}  // namespace ik
