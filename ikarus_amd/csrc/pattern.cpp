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

// elements per staging chunk (about 85 MB of FP64 element blocks)
constexpr int64_t kChunkElements = 32768;

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
  // Gather lists. Element e stores its 36 upper blocks (a <= b) at sb(a,b) = 8a - a(a-1)/2 + (b-a);
  // block (node_a, node_b) receives +block sb(a,b) when a <= b and +block(sb(b,a))^T otherwise.
  require((int64_t)ne * 72 < (int64_t)INT32_MAX, IK_EINVAL, "mesh too large for int32 gather codes");
  std::vector<int32_t> kcnt(h->nnzb + 1, 0);
  std::vector<int32_t> slots((size_t)ne * 64);
  for (int64_t e = 0; e < ne; ++e)
    for (int a = 0; a < 8; ++a) {
      const int32_t r = conn[e * 8 + a];
      const auto b = h->hcolidx.begin() + h->hrowptr[r], en = h->hcolidx.begin() + h->hrowptr[r + 1];
      for (int c = 0; c < 8; ++c) {
        const auto it = std::lower_bound(b, en, conn[e * 8 + c]);
        const int32_t p = (int32_t)(it - h->hcolidx.begin());
        slots[e * 64 + a * 8 + c] = p;
        kcnt[p + 1]++;
      }
    }
  for (int64_t p = 0; p < h->nnzb; ++p) kcnt[p + 1] += kcnt[p];
  std::vector<int32_t> klist(kcnt[h->nnzb]);
  {
    std::vector<int32_t> pos(kcnt.begin(), kcnt.end() - 1);
    for (int64_t e = 0; e < ne; ++e)  // increasing e: each list is ordered by element
      for (int a = 0; a < 8; ++a)
        for (int c = 0; c < 8; ++c) {
          const int lo = a < c ? a : c, hi = a < c ? c : a;
          const int sb = 8 * lo - (lo * (lo - 1)) / 2 + (hi - lo);
          const int32_t code = (int32_t)(((e * 36 + sb) << 1) | (a > c ? 1 : 0));
          klist[pos[slots[e * 64 + a * 8 + c]]++] = code;
        }
  }
  std::vector<int32_t> brow(h->nnzb);
  for (int64_t i = 0; i < nn; ++i)
    for (int32_t p = h->hrowptr[i]; p < h->hrowptr[i + 1]; ++p) brow[p] = (int32_t)i;
  // node incidences (e*8 + a), ordered by element
  std::vector<int32_t> ncnt(nn + 1, 0), nlist((size_t)ne * 8);
  for (int64_t i = 0; i < nn; ++i) ncnt[i + 1] = (int32_t)(cnt[i + 1] - cnt[0]);
  {
    std::vector<int32_t> pos(ncnt.begin(), ncnt.end() - 1);
    for (int64_t e = 0; e < ne; ++e)
      for (int a = 0; a < 8; ++a) nlist[pos[conn[e * 8 + a]]++] = (int32_t)(e * 8 + a);
  }
  // Element chunks: a block row is complete once its last incident element is staged. With the
  // prefix maximum of that index over rows (monotone by construction), rows [0, r) are complete
  // after all elements below pm[r-1], so each chunk's gather covers a contiguous row range right
  // behind the chunk while its staged blocks are still cached. Any numbering is correct; lexicographic
  // numberings keep the lag at about one row of nodes.
  std::vector<int64_t> last(nn, -1);
  for (int64_t i = 0; i < nn; ++i) {
    for (int64_t k = cnt[i]; k < cnt[i + 1]; ++k) last[i] = std::max(last[i], inc[k]);
    if (i > 0) last[i] = std::max(last[i], last[i - 1]);
  }
  const int64_t chunk = std::max<int64_t>(kChunkElements, 1);
  h->chunkElem.clear();
  h->chunkRow.clear();
  h->chunkElem.push_back(0);
  h->chunkRow.push_back(0);
  int64_t r = 0;
  for (int64_t e0 = 0; e0 < ne; e0 += chunk) {
    const int64_t e1 = std::min(ne, e0 + chunk);
    if (e1 == ne)
      r = nn;
    else
      while (r < nn && last[r] < e1) ++r;
    h->chunkElem.push_back(e1);
    h->chunkRow.push_back(r);
  }
  h->rowptr.upload(h->hrowptr.data(), h->hrowptr.size(), h->stream);
  h->colidx.upload(h->hcolidx.data(), h->hcolidx.size(), h->stream);
  h->diagBlk.upload(h->hdiag.data(), h->hdiag.size(), h->stream);
  h->blockRow.upload(brow.data(), brow.size(), h->stream);
  h->kcptr.upload(kcnt.data(), kcnt.size(), h->stream);
  h->kclist.upload(klist.data(), klist.size(), h->stream);
  h->ncptr.upload(ncnt.data(), ncnt.size(), h->stream);
  h->nclist.upload(nlist.data(), nlist.size(), h->stream);
  IK_HIP(hipStreamSynchronize(h->stream));
  h->patternBuilt = true;
}

}  // namespace ik
