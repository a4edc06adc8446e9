// pattern_sanitize.cpp — runs the threaded host pattern build (ikarus_amd/csrc/pattern.cpp) on the CPU under a
// sanitizer (built twice: AddressSanitizer + UndefinedBehaviorSanitizer, and ThreadSanitizer).
//
// The build's calls into the HIP runtime (device allocation, host-to-device copies, stream synchronisation) are
// replaced by host-memory test doubles defined in this file, so every phase of the build runs without a GPU and
// the "uploaded" arrays are plain host arrays. For each mesh the harness builds the pattern with 1, 3 and 8 host
// threads and requires every array to be byte-identical across thread counts.
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../ikarus_amd/csrc/ik_internal.hh"

extern "C" {
hipError_t hipMalloc(void** p, size_t n) {
  *p = std::malloc(n ? n : 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
  std::free(p);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
  std::memcpy(d, s, n);
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "host test double"; }
}

namespace {

struct Mesh {
  int64_t nn = 0, ne = 0;
  std::vector<double> X;
  std::vector<int32_t> conn;
  std::vector<uint8_t> fixed;
};

// structured n0 x n1 x n2 hex mesh (Dune corner order), bottom face clamped; optionally with a quarter of the
// elements dropped, the rest shuffled and the nodes relabelled
Mesh block(int n0, int n1, int n2, bool scramble, unsigned seed) {
  Mesh m;
  const int v0 = n0 + 1, v1 = n1 + 1, v2 = n2 + 1;
  m.nn = (int64_t)v0 * v1 * v2;
  m.X.resize(m.nn * 3);
  for (int k = 0; k < v2; ++k)
    for (int j = 0; j < v1; ++j)
      for (int i = 0; i < v0; ++i) {
        const int64_t v = i + v0 * (j + (int64_t)v1 * k);
        m.X[v * 3]      = i;
        m.X[v * 3 + 1]  = j;
        m.X[v * 3 + 2]  = k;
      }
  std::vector<std::array<int32_t, 8>> el;
  for (int k = 0; k < n2; ++k)
    for (int j = 0; j < n1; ++j)
      for (int i = 0; i < n0; ++i) {
        std::array<int32_t, 8> c;
        for (int a = 0; a < 8; ++a)
          c[a] = (i + (a & 1)) + v0 * ((j + ((a >> 1) & 1)) + v1 * (k + ((a >> 2) & 1)));
        el.push_back(c);
      }
  std::vector<int32_t> relabel(m.nn);
  for (int64_t v = 0; v < m.nn; ++v) relabel[v] = (int32_t)v;
  if (scramble) {
    std::mt19937 rng(seed);
    std::shuffle(el.begin(), el.end(), rng);
    el.resize(el.size() * 3 / 4);
    std::shuffle(relabel.begin(), relabel.end(), rng);
    std::vector<double> X2(m.X.size());
    for (int64_t v = 0; v < m.nn; ++v)
      for (int d = 0; d < 3; ++d) X2[relabel[v] * 3 + d] = m.X[v * 3 + d];
    m.X.swap(X2);
  }
  m.ne = (int64_t)el.size();
  for (auto& c : el)
    for (int a = 0; a < 8; ++a) m.conn.push_back(relabel[c[a]]);
  m.fixed.assign(m.nn * 3, 0);
  for (int64_t v = 0; v < m.nn; ++v)
    if (m.X[v * 3 + 2] == 0.0)
      for (int d = 0; d < 3; ++d) m.fixed[v * 3 + d] = 1;
  return m;
}

template <class T>
void put(std::string& out, const ik::DevBuf<T>& b) {
  const uint64_t n = b.p ? b.n : 0;
  out.append(reinterpret_cast<const char*>(&n), sizeof(n));
  if (n) out.append(reinterpret_cast<const char*>(b.p), n * sizeof(T));
}
template <class T, class A>
void put(std::string& out, const std::vector<T, A>& v) {
  const uint64_t n = v.size();
  out.append(reinterpret_cast<const char*>(&n), sizeof(n));
  if (n) out.append(reinterpret_cast<const char*>(v.data()), n * sizeof(T));
}

// every array the build produces, as bytes
std::string buildOnce(const Mesh& m, int eas, int threads) {
  setenv("IK_PATTERN_THREADS", std::to_string(threads).c_str(), 1);
  ik_handle h;
  h.nn   = m.nn;
  h.ne   = m.ne;
  h.ndof = m.nn * 3;
  h.neas = eas;
  h.hX   = m.X;
  h.hconn  = m.conn;
  h.hfixed = m.fixed;
  ik::buildPattern(&h);
  std::string out;
  put(out, h.hrowptr);
  put(out, h.hcolidx);
  put(out, h.hdiag);
  put(out, h.chunkQ);
  put(out, h.chunkElem);
  put(out, h.chunkRow);
  put(out, h.rowptr);
  put(out, h.colidx);
  put(out, h.diagBlk);
  put(out, h.hpos);
  put(out, h.sellStart);
  put(out, h.scol);
  put(out, h.upBlk);
  put(out, h.upT);
  put(out, h.upMask);
  put(out, h.bdest);
  put(out, h.segptr);
  put(out, h.ncptr);
  put(out, h.rdest);
  put(out, h.tileElem);
  put(out, h.tgptr);
  put(out, h.gdst);
  put(out, h.gpt);
  put(out, h.epos);
  put(out, h.gmsk);
  put(out, h.tmaxr);
  const int64_t sizes[4] = {h.nnzb, h.ntiles, h.nstage, h.nsell};
  out.append(reinterpret_cast<const char*>(sizes), sizeof(sizes));
  return out;
}

}  // namespace

int main() {
  struct Case {
    const char* name;
    Mesh m;
    int eas;
  };
  std::vector<Case> cases;
  // the build splits a phase over threads in chunks of at least 2048 nodes, elements or tiles: the meshes are
  // large enough that the node, element and tile phases all run on several threads
  cases.push_back({"block 48x48x32", block(48, 48, 32, false, 0), 0});
  cases.push_back({"block 32x32x20 EAS", block(32, 32, 20, false, 0), 21});
  cases.push_back({"scrambled 40x40x30", block(40, 40, 30, true, 7), 0});
  cases.push_back({"block 17x3x2", block(17, 3, 2, false, 0), 0});
  int bad = 0;
  for (auto& c : cases) {
    const std::string ref = buildOnce(c.m, c.eas, 1);
    for (int t : {3, 8}) {
      const std::string got = buildOnce(c.m, c.eas, t);
      if (got != ref) {
        std::printf("MISMATCH %s with %d threads\n", c.name, t);
        ++bad;
      }
    }
    std::printf("%s: %lld elements, %zu bytes of pattern arrays, threads 1/3/8 %s\n", c.name, (long long)c.m.ne,
                ref.size(), bad ? "differ" : "identical");
  }
  std::printf(bad ? "FAILED\n" : "OK\n");
  return bad ? 1 : 0;
}
