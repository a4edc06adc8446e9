// pattern.cpp — 3x3-block CSR sparsity pattern and element block-slot maps.
//
// Ikarus builds the scalar pattern from the union of element DOF cliques with Eigen setFromTriplets
// and keeps per element (3*8)^2 positions into the CSC value array (simpleassemblers.inl:207-222,
// 254-265; eigensparseaddon.hh:19-29). For FlatInterleaved DOFs (dof = 3*node + k, fehelper.hh:146-152)
// that scalar pattern is exactly the node-adjacency pattern expanded by 3x3 blocks, so the device keeps
// the block form: block row = node, sorted block columns, and per hex8 element 36 int32 staging slots
// (one per upper element block, pointing into the gather order of its K block). The scalar CSC is
// re-derived from it on demand.
#include <algorithm>
#include <cstdlib>

#include "ik_internal.hh"

namespace ik {

// elements per staging chunk; IKARUS_AMD_CHUNK_ELEMENTS overrides it. Splitting the 64^3 mesh into chunks
// whose gathers overlap the next chunk's element work measured slower than one launch of each (the two
// kernels contend for the same memory path), so the default takes meshes up to 1 M elements in one chunk.
constexpr int64_t kChunkElements = 1 << 20;

static int64_t chunkElements() {
  if (const char* v = std::getenv("IKARUS_AMD_CHUNK_ELEMENTS")) {
    const long long c = std::atoll(v);
    if (c >= 64) return (int64_t)c;
  }
  return kChunkElements;
}

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
  // Element e computes its 36 upper blocks (a <= b) sb(a,b) = 8a - a(a-1)/2 + (b-a). Block (node_a, node_b)
  // is an upper block of K when node_a <= node_b and receives the element block as it is; otherwise the
  // upper block (node_b, node_a) receives its transpose. Staging slots are grouped by upper block and
  // ordered by element inside each group, so the gather reads contiguous segments and sums them in a
  // fixed order.
  require((int64_t)ne * 36 < (int64_t)INT32_MAX / 2, IK_EINVAL, "mesh too large for int32 staging slots");
  std::vector<int32_t> slots((size_t)ne * 36);  // block index p of each element block
  std::vector<int32_t> kcnt(h->nnzb + 1, 0);
  for (int64_t e = 0; e < ne; ++e)
    for (int a = 0; a < 8; ++a)
      for (int c = a; c < 8; ++c) {
        const int32_t na = conn[e * 8 + a], nc = conn[e * 8 + c];
        const int32_t r = na < nc ? na : nc, col = na < nc ? nc : na;
        const auto b = h->hcolidx.begin() + h->hrowptr[r], en = h->hcolidx.begin() + h->hrowptr[r + 1];
        const int32_t p = (int32_t)(std::lower_bound(b, en, col) - h->hcolidx.begin());
        slots[e * 36 + 8 * a - (a * (a - 1)) / 2 + (c - a)] = p;
        kcnt[p + 1]++;
      }
  // upper blocks and their transposes
  h->huptr.assign(nn + 1, 0);
  for (int64_t i = 0; i < nn; ++i) h->huptr[i + 1] = h->huptr[i] + (h->hrowptr[i + 1] - h->hdiag[i]);
  std::vector<int32_t> up(h->huptr[nn]), upT(h->huptr[nn]);
  for (int64_t i = 0; i < nn; ++i)
    for (int32_t p = h->hdiag[i]; p < h->hrowptr[i + 1]; ++p) {
      const int32_t c = h->hcolidx[p];
      const auto b = h->hcolidx.begin() + h->hrowptr[c], en = h->hcolidx.begin() + h->hrowptr[c + 1];
      const int64_t q = h->huptr[i] + (p - h->hdiag[i]);
      up[q]           = p;
      upT[q]          = (int32_t)(std::lower_bound(b, en, (int32_t)i) - h->hcolidx.begin());
    }
  // segments in upper-block order, then each element block's slot (elements visited in increasing order)
  std::vector<int32_t> seg(h->huptr[nn] + 1, 0);
  for (int64_t q = 0; q < h->huptr[nn]; ++q) seg[q + 1] = seg[q] + kcnt[up[q] + 1];
  std::vector<int32_t> qOf(h->nnzb, -1);
  for (int64_t q = 0; q < h->huptr[nn]; ++q) qOf[up[q]] = (int32_t)q;
  std::vector<int32_t> dest((size_t)ne * 36);
  {
    std::vector<int32_t> pos(seg.begin(), seg.end() - 1);
    for (int64_t e = 0; e < ne; ++e)
      for (int a = 0; a < 8; ++a)
        for (int c = a; c < 8; ++c) {
          const int sb    = 8 * a - (a * (a - 1)) / 2 + (c - a);
          const int32_t q = qOf[slots[e * 36 + sb]];
          const bool tr   = conn[e * 8 + a] > conn[e * 8 + c];
          dest[e * 36 + sb] = (pos[q]++ << 1) | (tr ? 1 : 0);
        }
    for (int64_t q = 0; q < h->huptr[nn]; ++q)
      require(pos[q] == seg[q + 1], IK_EINVAL, "internal error: staging segments do not tile the slots");
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
  const int64_t chunk = chunkElements();
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
  std::vector<uint16_t> mask(up.size(), 0);
  for (int64_t i = 0; i < nn; ++i)
    for (int32_t p = h->hdiag[i]; p < h->hrowptr[i + 1]; ++p) {
      const int64_t c = h->hcolidx[p];
      uint16_t m      = 0;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l)
          if (h->hfixed[3 * i + k] | h->hfixed[3 * c + l]) {
            m |= (uint16_t)(1u << (3 * k + l));
            if (3 * i + k == 3 * c + l) m |= (uint16_t)(1u << (9 + k));
          }
      mask[h->huptr[i] + (p - h->hdiag[i])] = m;
    }
  h->upMask.upload(mask.data(), mask.size(), h->stream);
  h->upBlk.upload(up.data(), up.size(), h->stream);
  h->upT.upload(upT.data(), upT.size(), h->stream);
  h->bdest.upload(dest.data(), dest.size(), h->stream);
  h->segptr.upload(seg.data(), seg.size(), h->stream);
  // residual staging slot of every element corner: its position in the node's incidence list
  std::vector<int32_t> rdest((size_t)ne * 8);
  for (size_t j = 0; j < nlist.size(); ++j) rdest[nlist[j]] = (int32_t)j;
  h->ncptr.upload(ncnt.data(), ncnt.size(), h->stream);
  h->rdest.upload(rdest.data(), rdest.size(), h->stream);
  IK_HIP(hipStreamSynchronize(h->stream));
  h->patternBuilt = true;
}

}  // namespace ik
