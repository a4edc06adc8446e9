// host_parallel.hh — the host threads of the pattern and coarse-space builds.
#pragma once

#include <algorithm>
#include <cstdlib>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

#include "ik_internal.hh"

namespace ik {

// Host threads of the pattern and coarse-space builds: ranges of nodes, elements, tiles or aggregates whose
// outputs sit at positions fixed
// before the threads start (variable-length outputs go to per-chunk buffers placed by a prefix sum after
// the join), so every array is the same for any thread count. IK_PATTERN_THREADS sets the count (1:
// serial); the default is the host's hardware threads, at most 16.
inline int patternThreads() {
  if (const char* v = std::getenv("IK_PATTERN_THREADS")) {
    const int t = std::atoi(v);
    if (t >= 1) return std::min(t, 64);
  }
  const unsigned hc = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(hc, 16u));
}
// chunks of at least `grain` items (2048 nodes, elements or tiles; 32 slices of 64 rows)
inline int64_t chunkCount(int64_t n, int64_t grain = 2048) {
  return std::min<int64_t>(patternThreads(), std::max<int64_t>(1, n / grain));
}

// f(lo, hi, chunk) over chunk = 0 .. chunkCount(n) - 1; the first exception thrown is rethrown here
template <class F>
void parallelChunks(int64_t n, F&& f, int64_t grain = 2048) {
  const int64_t T = chunkCount(n, grain);
  if (T <= 1) {
    f(int64_t(0), n, int64_t(0));
    return;
  }
  std::vector<std::thread> th;
  std::exception_ptr err;
  std::mutex m;
  th.reserve(T);
  for (int64_t c = 0; c < T; ++c)
    th.emplace_back([&, c] {
      try {
        f(n * c / T, n * (c + 1) / T, c);
      } catch (...) {
        std::lock_guard<std::mutex> g(m);
        if (!err) err = std::current_exception();
      }
    });
  for (auto& x : th) x.join();
  if (err) std::rethrow_exception(err);
}

// out[i] = sum of w(k) over k < i, for i = 0 .. n (out has n + 1 entries): per-chunk sums, their prefix, then
// each chunk's running sum; the same values for any thread count (integer sums)
template <class W>
void parallelPrefix(int64_t n, W&& w, HostVec<int64_t>& out) {
  out.clear();
  out.resize(n + 1);
  const int64_t T = chunkCount(n);
  std::vector<int64_t> part(T + 1, 0);
  parallelChunks(n, [&](int64_t i0, int64_t i1, int64_t c) {
    int64_t s = 0;
    for (int64_t i = i0; i < i1; ++i) s += w(i);
    part[c + 1] = s;
  });
  for (int64_t c = 0; c < T; ++c) part[c + 1] += part[c];
  parallelChunks(n, [&](int64_t i0, int64_t i1, int64_t c) {
    int64_t s = part[c];
    for (int64_t i = i0; i < i1; ++i) {
      out[i] = s;
      s += w(i);
    }
    if (i1 == n) out[n] = s;
  });
}

// n copies of v, written by the build's threads
template <class T>
HostVec<T> filled(int64_t n, T v) {
  HostVec<T> b(n);
  parallelChunks(n, [&](int64_t i0, int64_t i1, int64_t) { std::fill(b.data() + i0, b.data() + i1, v); });
  return b;
}

}  // namespace ik
