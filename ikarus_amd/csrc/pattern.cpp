// pattern.cpp — 3x3-block CSR sparsity pattern and element block-slot maps.
//
// Ikarus builds the scalar pattern from the union of element DOF cliques with Eigen setFromTriplets
// and keeps per element (3*8)^2 positions into the CSC value array (simpleassemblers.inl:207-222,
// 254-265; eigensparseaddon.hh:19-29). For FlatInterleaved DOFs (dof = 3*node + k, fehelper.hh:146-152)
// that scalar pattern is exactly the node-adjacency pattern expanded by 3x3 blocks, so the device keeps
// the block form: block row = node, sorted block columns, and 64 int32 block slots per hex8 element
// (slot[a*8+b] = position of block (node_a, node_b)). The scalar CSC is re-derived from it on demand.
#include <algorithm>

#include "ik_internal.hh"

namespace ik {

void buildPattern(ik_handle* h) {
  const int64_t nn = h->nn, ne = h->ne;
  const int32_t* conn = h->hconn.data();
  // node -> incident elements (CSR)
  std::vector<int64_t> cnt(nn + 1, 0);
  for (int64_t e = 0; e < ne; ++e)
    for (int a = 0; a < 8; ++a) cnt[conn[e * 8 + a] + 1]++;
  for (int64_t i = 0; i < nn; ++i) cnt[i + 1] += cnt[i];
  std::vector<int64_t> inc(cnt[nn]);
  {
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t e = 0; e < ne; ++e)
      for (int a = 0; a < 8; ++a) inc[pos[conn[e * 8 + a]]++] = e;
  }
  // rows: union of the nodes of incident elements
  h->hrowptr.assign(nn + 1, 0);
  std::vector<int32_t> cols;
  cols.reserve((size_t)nn * 27);
  std::vector<int32_t> buf;
  for (int64_t i = 0; i < nn; ++i) {
    buf.clear();
    for (int64_t k = cnt[i]; k < cnt[i + 1]; ++k) {
      const int64_t e = inc[k];
      for (int b = 0; b < 8; ++b) buf.push_back(conn[e * 8 + b]);
    }
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
    cols.insert(cols.end(), buf.begin(), buf.end());
    h->hrowptr[i + 1] = (int32_t)cols.size();
  }
  require(cols.size() < (size_t)INT32_MAX, IK_EINVAL, "pattern too large for int32 block indices");
  h->hcolidx.swap(cols);
  h->nnzb = (int64_t)h->hcolidx.size();
  h->hdiag.assign(nn, -1);
  for (int64_t i = 0; i < nn; ++i) {
    const auto b = h->hcolidx.begin() + h->hrowptr[i], en = h->hcolidx.begin() + h->hrowptr[i + 1];
    h->hdiag[i] = (int32_t)(std::lower_bound(b, en, (int32_t)i) - h->hcolidx.begin());
  }
  std::vector<int32_t> slots((size_t)ne * 64);
  for (int64_t e = 0; e < ne; ++e)
    for (int a = 0; a < 8; ++a) {
      const int32_t r = conn[e * 8 + a];
      const auto b = h->hcolidx.begin() + h->hrowptr[r], en = h->hcolidx.begin() + h->hrowptr[r + 1];
      for (int c = 0; c < 8; ++c) {
        const auto it = std::lower_bound(b, en, conn[e * 8 + c]);
        slots[e * 64 + a * 8 + c] = (int32_t)(it - h->hcolidx.begin());
      }
    }
  h->rowptr.upload(h->hrowptr.data(), h->hrowptr.size(), h->stream);
  h->colidx.upload(h->hcolidx.data(), h->hcolidx.size(), h->stream);
  h->diagBlk.upload(h->hdiag.data(), h->hdiag.size(), h->stream);
  h->bslot.upload(slots.data(), slots.size(), h->stream);
  IK_HIP(hipStreamSynchronize(h->stream));
  h->patternBuilt = true;
}

}  // namespace ik
