// pattern.cpp — 3x3-block sparsity pattern, its sliced-ELL storage slots and element block-slot maps.
//
// Ikarus builds the scalar pattern from the union of element DOF cliques with Eigen setFromTriplets
// and keeps per element (3*8)^2 positions into the CSC value array (simpleassemblers.inl:207-222,
// 254-265; eigensparseaddon.hh:19-29). For FlatInterleaved DOFs (dof = 3*node + k, fehelper.hh:146-152)
// that scalar pattern is exactly the node-adjacency pattern expanded by 3x3 blocks, so the device keeps
// the block form: block row = node, sorted block columns (CSR order), each block stored at a sliced-ELL
// slot, and the maps from element blocks to staging slots in the gather order of their K blocks (per
// element for EAS, per tile group for the plain element). The scalar CSC is re-derived from it on demand.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <thread>

#include "host_parallel.hh"
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

// the low 21 bits of v spread to every third bit (bit i -> bit 3i)
static uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0x1FFFFFu;
  x          = (x | x << 32) & 0x1F00000000FFFFull;
  x          = (x | x << 16) & 0x1F0000FF0000FFull;
  x          = (x | x << 8) & 0x100F00F00F00F00Full;
  x          = (x | x << 4) & 0x10C30C30C30C30C3ull;
  x          = (x | x << 2) & 0x1249249249249249ull;
  return x;
}

// The mesh lattice: node bounds and, per axis, the median element extent, an element's extent along axis d
// being the largest |component d| of its three mean edge vectors (the average of the four edges along each
// reference axis), which node jitter leaves unbiased. On a structured grid the lattice is the grid itself.
static void meshLattice(ik_handle* h) {
  const int64_t ne = h->ne, nn = h->nn;
  const double* X     = h->hX.data();
  const int32_t* conn = h->hconn.data();
  const int64_t T     = chunkCount(nn);
  std::vector<double> plo((size_t)T * 3, 0.0), phi((size_t)T * 3, 0.0);
  parallelChunks(nn, [&](int64_t n0, int64_t n1, int64_t c) {
    for (int d = 0; d < 3; ++d) {
      double lo = X[n0 * 3 + d], hi = lo;
      for (int64_t n = n0 + 1; n < n1; ++n) {
        lo = std::min(lo, X[n * 3 + d]);
        hi = std::max(hi, X[n * 3 + d]);
      }
      plo[c * 3 + d] = lo;
      phi[c * 3 + d] = hi;
    }
  });
  HostVec<double> ext(ne);
  for (int d = 0; d < 3; ++d) {
    h->latLo[d] = plo[d];
    h->latHi[d] = phi[d];
    for (int64_t c = 1; c < T; ++c) {
      h->latLo[d] = std::min(h->latLo[d], plo[c * 3 + d]);
      h->latHi[d] = std::max(h->latHi[d], phi[c * 3 + d]);
    }
    parallelChunks(ne, [&](int64_t e0, int64_t e1, int64_t) {
      for (int64_t e = e0; e < e1; ++e) {
        double best = 0.0;
        for (int m = 0; m < 3; ++m) {
          double v = 0.0;  // 4 x component d of the mean edge vector along reference axis m
          for (int a = 0; a < 8; ++a)
            if (!((a >> m) & 1)) v += X[conn[e * 8 + (a | (1 << m))] * 3 + d] - X[conn[e * 8 + a] * 3 + d];
          best = std::max(best, std::fabs(0.25 * v));
        }
        ext[e] = best;
      }
    });
    std::nth_element(ext.begin(), ext.begin() + ne / 2, ext.end());
    h->latStep[d] = ext[ne / 2] > 0.0 ? ext[ne / 2] : 1.0;
  }
}

// Elements in Morton (Z) order of their centroids on the mesh lattice, ties by element index. On a
// structured grid 16 consecutive elements then form a 4x2x2 box (45 nodes) whose 576 upper element blocks
// fall on 341 distinct K blocks; on other meshes the order still keeps each tile compact.
static HostVec<int32_t> mortonOrder(const ik_handle* h) {
  const int64_t ne = h->ne;
  const double* X     = h->hX.data();
  const int32_t* conn = h->hconn.data();
  const double* lo = h->latLo;
  const double* step = h->latStep;
  HostVec<std::pair<uint64_t, int32_t>> key(ne);
  parallelChunks(ne, [&](int64_t e0, int64_t e1, int64_t) {
    for (int64_t e = e0; e < e1; ++e) {
      uint64_t k = 0;
      for (int d = 0; d < 3; ++d) {
        double c = 0.0;
        for (int a = 0; a < 8; ++a) c += X[conn[e * 8 + a] * 3 + d];
        const double q = std::floor((0.125 * c - lo[d]) / step[d]);
        const uint32_t qi = q <= 0.0 ? 0u : (q >= 2097151.0 ? 2097151u : (uint32_t)q);
        k |= spread3(qi) << d;
      }
      key[e] = {k, (int32_t)e};
    }
  });
  // sorted runs per thread, then merged pairwise (keys are distinct: the element index breaks ties)
  {
    const int64_t T = chunkCount(ne);
    std::vector<int64_t> b(T + 1);
    for (int64_t c = 0; c <= T; ++c) b[c] = ne * c / T;
    parallelChunks(ne, [&](int64_t e0, int64_t e1, int64_t) { std::sort(key.begin() + e0, key.begin() + e1); });
    for (int64_t w = 1; w < T; w *= 2) {
      std::vector<std::thread> th;
      for (int64_t c = 0; c + w < T; c += 2 * w)
        th.emplace_back([&, c] {
          std::inplace_merge(key.begin() + b[c], key.begin() + b[c + w], key.begin() + b[std::min(T, c + 2 * w)]);
        });
      for (auto& x : th) x.join();
    }
  }
  HostVec<int32_t> order(ne);
  parallelChunks(ne, [&](int64_t i0, int64_t i1, int64_t) {
    for (int64_t i = i0; i < i1; ++i) order[i] = key[i].second;
  });
  return order;
}

void buildPattern(ik_handle* h) {
  const int64_t nn = h->nn, ne = h->ne;
  const int32_t* conn = h->hconn.data();
  meshLattice(h);
  // node -> incidences e*8 + a (CSR, ordered by element): nlist; the incident elements are nlist / 8
  std::vector<int64_t> cnt(nn + 1, 0);
  for (int64_t e = 0; e < ne; ++e)
    for (int a = 0; a < 8; ++a) cnt[conn[e * 8 + a] + 1]++;
  for (int64_t i = 0; i < nn; ++i) cnt[i + 1] += cnt[i];
  require(cnt[nn] < (int64_t)INT32_MAX, IK_EINVAL, "mesh too large for int32 incidence lists");
  HostVec<int32_t> nlist(cnt[nn]);
  {
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t e = 0; e < ne; ++e)
      for (int a = 0; a < 8; ++a) nlist[pos[conn[e * 8 + a]]++] = (int32_t)(e * 8 + a);
  }
  auto inc = [&](int64_t k) { return (int64_t)(nlist[k] >> 3); };
  // rows: union of the nodes of incident elements (per chunk of rows, then placed by the row counts)
  h->hrowptr.assign(nn + 1, 0);
  {
    std::vector<std::vector<int32_t>> part(chunkCount(nn));
    HostVec<int64_t> rc(nn);
    parallelChunks(nn, [&](int64_t i0, int64_t i1, int64_t c) {
      std::vector<int32_t>& cols = part[c];
      cols.reserve((size_t)(i1 - i0) * 27);
      std::vector<int32_t> buf;
      for (int64_t i = i0; i < i1; ++i) {
        buf.clear();
        for (int64_t k = cnt[i]; k < cnt[i + 1]; ++k) {
          const int64_t e = inc(k);
          for (int b = 0; b < 8; ++b) buf.push_back(conn[e * 8 + b]);
        }
        std::sort(buf.begin(), buf.end());
        buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
        cols.insert(cols.end(), buf.begin(), buf.end());
        rc[i] = (int64_t)buf.size();
      }
    });
    int64_t tot = 0;
    for (int64_t i = 0; i < nn; ++i) {
      tot += rc[i];
      require(tot < (int64_t)INT32_MAX, IK_EINVAL, "pattern too large for int32 block indices");
      h->hrowptr[i + 1] = (int32_t)tot;
    }
    h->hcolidx.clear();
    h->hcolidx.resize(tot);
    parallelChunks(nn, [&](int64_t i0, int64_t, int64_t c) {
      std::copy(part[c].begin(), part[c].end(), h->hcolidx.begin() + h->hrowptr[i0]);
    });
  }
  h->nnzb = (int64_t)h->hcolidx.size();
  h->hdiag.assign(nn, -1);
  parallelChunks(nn, [&](int64_t i0, int64_t i1, int64_t) {
    for (int64_t i = i0; i < i1; ++i) {
      const auto b = h->hcolidx.begin() + h->hrowptr[i], en = h->hcolidx.begin() + h->hrowptr[i + 1];
      h->hdiag[i] = (int32_t)(std::lower_bound(b, en, (int32_t)i) - h->hcolidx.begin());
    }
  });
  // Sliced-ELL slots of K: slice s (rows [64 s, 64 s + 64)) is as wide as its longest row; block j of row i
  // sits at slot sstart[s] + 64 j + (i mod 64). The last slice is padded to 64 rows too (rows past nn own only
  // padding slots, which the kernels never read).
  const int64_t nslices = (nn + kSellRows - 1) / kSellRows;
  std::vector<int32_t> sstart(nslices + 1, 0);
  {
    int64_t tot = 0;
    for (int64_t s = 0; s < nslices; ++s) {
      int32_t w = 0;
      for (int64_t i = s * kSellRows; i < std::min(nn, (s + 1) * kSellRows); ++i)
        w = std::max(w, h->hrowptr[i + 1] - h->hrowptr[i]);
      tot += (int64_t)w * kSellRows;
      require(tot < (int64_t)INT32_MAX, IK_EINVAL, "pattern too large for int32 block slots");
      sstart[s + 1] = (int32_t)tot;
    }
  }
  h->nsell = sstart[nslices];
  h->hpos.clear();
  h->hpos.resize(h->nnzb);  // every block gets its slot below
  HostVec<int32_t> scol(h->nsell);
  parallelChunks(nslices, [&](int64_t s0, int64_t s1, int64_t) {
    for (int64_t s = s0; s < s1; ++s) {
      const int32_t w = (sstart[s + 1] - sstart[s]) / kSellRows;
      for (int r = 0; r < kSellRows; ++r) {
        const int64_t i = s * kSellRows + r;
        const int32_t len = i < nn ? h->hrowptr[i + 1] - h->hrowptr[i] : 0;
        for (int32_t j = 0; j < w; ++j) {
          const int32_t slot = sstart[s] + j * kSellRows + r;
          if (j < len) {
            h->hpos[h->hrowptr[i] + j] = slot;
            scol[slot]                 = h->hcolidx[h->hrowptr[i] + j];
          } else {
            scol[slot] = (int32_t)std::min(i, nn - 1);
          }
        }
      }
    }
  }, 32);
  {
    // every block in its row's slice, at a slot of its own holding its column (slices are disjoint slot ranges,
    // so each thread marks only its own slices' slots)
    HostVec<uint8_t> seen = filled<uint8_t>(h->nsell, 0);
    parallelChunks(nslices, [&](int64_t s0, int64_t s1, int64_t) {
      for (int64_t i = s0 * kSellRows; i < std::min(nn, s1 * kSellRows); ++i)
        for (int32_t b = h->hrowptr[i]; b < h->hrowptr[i + 1]; ++b) {
          const int32_t sl = h->hpos[b];
          require(sl >= sstart[i / kSellRows] && sl < sstart[i / kSellRows + 1] && !seen[sl] &&
                      scol[sl] == h->hcolidx[b],
                  IK_EINVAL, "internal error: bad sliced-ELL slot");
          seen[sl] = 1;
        }
    }, 32);
  }
  h->sellStart.upload(sstart.data(), sstart.size(), h->stream);
  h->scol.upload(scol.data(), scol.size(), h->stream);
  // Element e computes its 36 upper blocks (a <= b) sb(a,b) = 8a - a(a-1)/2 + (b-a). Block (node_a, node_b)
  // is an upper block of K when node_a <= node_b and receives the element block as it is; otherwise the
  // upper block (node_b, node_a) receives its transpose. Staging slots are grouped by upper block and
  // ordered by staging unit (tile or element) inside each group, so the gather reads contiguous segments
  // and sums them in a fixed order.
  require((int64_t)ne * 36 < (int64_t)INT32_MAX / 2, IK_EINVAL, "mesh too large for int32 staging slots");
  HostVec<int32_t> slots((size_t)ne * 36);  // block index p of each element block
  parallelChunks(ne, [&](int64_t e0, int64_t e1, int64_t) {
    for (int64_t e = e0; e < e1; ++e)
      for (int a = 0; a < 8; ++a)
        for (int c = a; c < 8; ++c) {
          const int32_t na = conn[e * 8 + a], nc = conn[e * 8 + c];
          const int32_t r = na < nc ? na : nc, col = na < nc ? nc : na;
          const auto b = h->hcolidx.begin() + h->hrowptr[r], en = h->hcolidx.begin() + h->hrowptr[r + 1];
          slots[e * 36 + 8 * a - (a * (a - 1)) / 2 + (c - a)] =
              (int32_t)(std::lower_bound(b, en, col) - h->hcolidx.begin());
        }
  });
  // upper blocks and their transposes
  std::vector<int64_t> uptr(nn + 1, 0);  // first upper block of each row
  for (int64_t i = 0; i < nn; ++i) uptr[i + 1] = uptr[i] + (h->hrowptr[i + 1] - h->hdiag[i]);
  HostVec<int32_t> up(uptr[nn]), upT(uptr[nn]);
  parallelChunks(nn, [&](int64_t i0, int64_t i1, int64_t) {
    for (int64_t i = i0; i < i1; ++i)
      for (int32_t p = h->hdiag[i]; p < h->hrowptr[i + 1]; ++p) {
        const int32_t c = h->hcolidx[p];
        const auto b = h->hcolidx.begin() + h->hrowptr[c], en = h->hcolidx.begin() + h->hrowptr[c + 1];
        const int64_t q = uptr[i] + (p - h->hdiag[i]);
        up[q]           = p;
        upT[q]          = (int32_t)(std::lower_bound(b, en, (int32_t)i) - h->hcolidx.begin());
      }
  });
  const int64_t nup = uptr[nn];
  HostVec<int32_t> qOf = filled<int32_t>(h->nnzb, -1);
  parallelChunks(nup, [&](int64_t q0, int64_t q1, int64_t) {
    for (int64_t q = q0; q < q1; ++q) qOf[up[q]] = (int32_t)q;
  });
  // FULL Dirichlet treatment per upper block: bit 3k+l zeroes entry (k,l), bit 9+k sets a unit diagonal
  HostVec<uint16_t> mask(nup);
  parallelChunks(nn, [&](int64_t i0, int64_t i1, int64_t) {
    for (int64_t i = i0; i < i1; ++i)
      for (int32_t p = h->hdiag[i]; p < h->hrowptr[i + 1]; ++p) {
        const int64_t c = h->hcolidx[p];
        uint16_t m      = 0;
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l)
            if (h->hfixed[3 * i + k] | h->hfixed[3 * c + l]) {
              m |= (uint16_t)(1u << (3 * k + l));
              if (3 * i + k == 3 * c + l) m |= (uint16_t)(1u << (9 + k));
            }
        mask[uptr[i] + (p - h->hdiag[i])] = m;
      }
  });
  // upper blocks the gather assembles (the rest are written directly)
  HostVec<uint8_t> gathered = filled<uint8_t>(nup, 1);
  HostVec<int64_t> unitOf(ne);  // staging unit of each element: its tile (tiled) or itself
  h->tiled = h->neas == 0;
  int64_t ntiles = 0;
  HostVec<int32_t> tileElem;
  if (h->tiled) {
    // plain element: tiles of kTileElems, consecutive runs of the Morton order (compact boxes), then sorted by
    // their lowest node, so tiles running at the same time write neighbouring parts of the segments.
    const HostVec<int32_t> order = mortonOrder(h);
    ntiles = (ne + kTileElems - 1) / kTileElems;
    std::vector<std::pair<int32_t, int32_t>> tkey(ntiles);  // (lowest node, Morton tile)
    parallelChunks(ntiles, [&](int64_t m0, int64_t m1, int64_t) {
      for (int64_t m = m0; m < m1; ++m) {
        int32_t lo = INT32_MAX;
        for (int64_t i = m * kTileElems; i < std::min(ne, (m + 1) * kTileElems); ++i)
          for (int a = 0; a < 8; ++a) lo = std::min(lo, conn[(int64_t)order[i] * 8 + a]);
        tkey[m] = {lo, (int32_t)m};
      }
    });
    std::sort(tkey.begin(), tkey.end());
    tileElem = filled<int32_t>(ntiles * kTileElems, -1);
    parallelChunks(ntiles, [&](int64_t t0, int64_t t1, int64_t) {
      for (int64_t t = t0; t < t1; ++t) {
        const int64_t m = tkey[t].second;
        for (int64_t i = m * kTileElems; i < std::min(ne, (m + 1) * kTileElems); ++i) {
          tileElem[t * kTileElems + (i - m * kTileElems)] = order[i];
          unitOf[order[i]] = t;
        }
      }
    });
  } else {
    for (int64_t e = 0; e < ne; ++e) unitOf[e] = e;
  }
  const int64_t nunits = h->tiled ? ntiles : ne;
  // Staging chunks (ranges of staging units: tiles, or elements for EAS): a block row is complete once the
  // last unit touching it is staged. With the prefix maximum of that unit index over rows (monotone by
  // construction), rows [0, r) are complete after all units below pm[r-1], so each chunk's gather covers a
  // contiguous row range behind the chunk. Any numbering is correct.
  {
    std::vector<int64_t> last(nn, -1);
    for (int64_t i = 0; i < nn; ++i) {
      for (int64_t k = cnt[i]; k < cnt[i + 1]; ++k) last[i] = std::max(last[i], unitOf[inc(k)]);
      if (i > 0) last[i] = std::max(last[i], last[i - 1]);
    }
    const int64_t chunk = h->tiled ? (chunkElements() + kTileElems - 1) / kTileElems : chunkElements();
    h->chunkElem.assign(1, 0);
    h->chunkRow.assign(1, 0);
    int64_t r = 0;
    for (int64_t u0 = 0; u0 < nunits; u0 += chunk) {
      const int64_t u1 = std::min(nunits, u0 + chunk);
      if (u1 == nunits)
        r = nn;
      else
        while (r < nn && last[r] < u1) ++r;
      h->chunkElem.push_back(u1);
      h->chunkRow.push_back(r);
    }
  }
  const int64_t nchunks = (int64_t)h->chunkElem.size() - 1;
  // Gather order of the upper blocks: chunk by chunk (the rows each chunk completes), inside a chunk by storage
  // slot. For a fixed j the rows of a slice have adjacent slots, so the gather's consecutive lanes, and a tile's
  // consecutive groups, write neighbouring slots. Staging segments follow this order; qrank[q] = position of q.
  std::vector<int32_t> qord;
  HostVec<int32_t> qrank = filled<int32_t>(nup, -1);
  std::vector<int64_t> qchunk(nchunks + 1, 0);  // qord[qchunk[c] .. qchunk[c+1]) = upper blocks of chunk c
  {
    HostVec<int32_t> qAtSlot = filled<int32_t>(h->nsell, -1);
    parallelChunks(nup, [&](int64_t q0, int64_t q1, int64_t) {
      for (int64_t q = q0; q < q1; ++q) qAtSlot[h->hpos[up[q]]] = (int32_t)q;
    });
    qord.reserve(nup);
    for (int64_t c = 0; c < nchunks; ++c) {
      const int64_t r0 = h->chunkRow[c], r1 = h->chunkRow[c + 1];
      if (r1 > r0) {
        // the chunk's slices in ranges per thread, each collecting its upper blocks in slot order, then joined
        const int64_t sl0 = r0 / kSellRows, nsl = (r1 - 1) / kSellRows + 1 - sl0;
        std::vector<std::vector<int32_t>> part(chunkCount(nsl, 32));
        parallelChunks(nsl, [&](int64_t a0, int64_t a1, int64_t ch) {
          for (int64_t sl = sl0 + a0; sl < sl0 + a1; ++sl)
            for (int32_t slot = sstart[sl]; slot < sstart[sl + 1]; ++slot) {
              const int32_t q   = qAtSlot[slot];
              const int64_t row = sl * kSellRows + (slot - sstart[sl]) % kSellRows;
              if (q >= 0 && row >= r0 && row < r1) part[ch].push_back(q);
            }
        }, 32);
        for (const auto& v : part) qord.insert(qord.end(), v.begin(), v.end());
      }
      qchunk[c + 1] = (int64_t)qord.size();
    }
    parallelChunks((int64_t)qord.size(), [&](int64_t k0, int64_t k1, int64_t) {
      for (int64_t k = k0; k < k1; ++k) qrank[qord[k]] = (int32_t)k;
    });
    require((int64_t)qord.size() == nup, IK_EINVAL, "internal error: gather order misses upper blocks");
  }
  // staging segment of upper block q: [segBeg[q], segBeg[q] + segCnt[q]), segments laid out in gather order
  HostVec<int32_t> segBeg(nup), segCnt(nup);  // both written for every upper block below
  auto layoutSegments = [&]() {
    HostVec<int64_t> beg;
    parallelPrefix((int64_t)qord.size(), [&](int64_t k) { return (int64_t)segCnt[qord[k]]; }, beg);
    const int64_t tot = beg.back();
    require(tot < (int64_t)INT32_MAX, IK_EINVAL, "mesh too large for int32 staging slots");
    parallelChunks((int64_t)qord.size(), [&](int64_t k0, int64_t k1, int64_t) {
      for (int64_t k = k0; k < k1; ++k) segBeg[qord[k]] = (int32_t)beg[k];
    });
    return tot;
  };
  if (!h->tiled) {
    // EAS: one slot per element block, elements in increasing order inside each segment
    std::vector<int32_t> kcnt(h->nnzb, 0);  // element blocks per K block
    for (const int32_t p : slots) kcnt[p]++;
    for (int64_t q = 0; q < nup; ++q) segCnt[q] = kcnt[up[q]];
    h->nstage = layoutSegments();
    std::vector<int32_t> dest((size_t)ne * 36);
    std::vector<int32_t> pos(segBeg.begin(), segBeg.end());
    for (int64_t e = 0; e < ne; ++e)
      for (int a = 0; a < 8; ++a)
        for (int c = a; c < 8; ++c) {
          const int sb    = 8 * a - (a * (a - 1)) / 2 + (c - a);
          const int32_t q = qOf[slots[e * 36 + sb]];
          const bool tr   = conn[e * 8 + a] > conn[e * 8 + c];
          dest[e * 36 + sb] = (pos[q]++ << 1) | (tr ? 1 : 0);
        }
    for (int64_t q = 0; q < nup; ++q)
      require(pos[q] == segBeg[q] + segCnt[q], IK_EINVAL, "internal error: staging segments do not tile the slots");
    h->bdest.upload(dest.data(), dest.size(), h->stream);
  } else {
    // per tile: the entries j = 36 s + sb that fall on one upper block form a group; groups ordered by the gather
    // order of their upper blocks, i.e. by staging slot and by K slot; each entry's rank = its position in its
    // group by increasing j (the order the tile sums them).
    // epos[t*576 + j] = group | rank << 10 | transposed << 15; padding entries target spare accumulators
    // ng + (j - 36 nvalid), rank 0, which the tile never stores. tmaxr[t] = the largest group size.
    // Each chunk of tiles lists its groups' upper blocks in its own buffer; tgptr places them.
    HostVec<uint16_t> epos((size_t)ntiles * kTileEntries);
    std::vector<uint8_t> tmaxr(ntiles);
    std::vector<int32_t> tgptr(ntiles + 1, 0);
    HostVec<int32_t> gq;
    std::vector<int64_t> tng(ntiles);
    // a tile's upper blocks -> group: open-addressing table of 2048 slots (a tile touches at most 576 upper blocks),
    // a slot valid when its stamp is the current tile
    constexpr int kHash = 2048;
    std::vector<std::vector<int32_t>> gpart(chunkCount(ntiles, 256));
    parallelChunks(ntiles, [&](int64_t t0, int64_t t1, int64_t ch) {
      std::vector<int32_t> hkey(kHash), hval(kHash), touched, entQ(kTileEntries), fill(kTileEntries);
      std::vector<int64_t> hstamp(kHash, -1);
      std::vector<uint16_t> entTr(kTileEntries);
      std::vector<int32_t>& out = gpart[ch];
      auto slotOf = [&](int32_t q, int64_t t) {  // the table slot of q in tile t (free or holding q)
        uint32_t k = ((uint32_t)q * 2654435761u) & (kHash - 1);
        while (hstamp[k] == t && hkey[k] != q) k = (k + 1) & (kHash - 1);
        return k;
      };
      for (int64_t t = t0; t < t1; ++t) {
        touched.clear();
        int nvalid = 0;
        for (int s = 0; s < kTileElems; ++s) {
          const int32_t e = tileElem[t * kTileElems + s];
          if (e < 0) continue;
          ++nvalid;
          for (int a = 0; a < 8; ++a)
            for (int c = a; c < 8; ++c) {
              const int sb    = 8 * a - (a * (a - 1)) / 2 + (c - a);
              const int j     = 36 * s + sb;
              const int32_t q = qOf[slots[(int64_t)e * 36 + sb]];
              entQ[j]         = q;
              entTr[j]        = conn[(int64_t)e * 8 + a] > conn[(int64_t)e * 8 + c] ? 0x8000u : 0u;
              const uint32_t k = slotOf(q, t);
              if (hstamp[k] != t) {
                hstamp[k] = t;
                hkey[k]   = q;
                touched.push_back(q);
              }
            }
        }
        // gather order: sort the ranks, then map back to upper blocks
        for (auto& q : touched) q = qrank[q];
        std::sort(touched.begin(), touched.end());
        for (auto& q : touched) q = qord[q];
        const int ng = (int)touched.size();
        for (int g = 0; g < ng; ++g) hval[slotOf(touched[g], t)] = g;
        std::fill(fill.begin(), fill.begin() + ng, 0);
        uint16_t* ep = epos.data() + t * kTileEntries;
        int maxr     = 0;
        for (int j = 0; j < kTileEntries; ++j) {
          if (j < 36 * nvalid) {
            const int g  = hval[slotOf(entQ[j], t)];
            const int rk = fill[g]++;
            require(rk < 16, IK_EINVAL, "more than 16 element blocks of one tile on one K block");
            maxr  = std::max(maxr, rk + 1);
            ep[j] = (uint16_t)(g | (rk << 10) | entTr[j]);
          } else {
            ep[j] = (uint16_t)(ng + (j - 36 * nvalid));
          }
        }
        tmaxr[t] = (uint8_t)maxr;
        out.insert(out.end(), touched.begin(), touched.end());
        tng[t] = ng;
      }
    }, 256);
    {
      int64_t tot = 0;
      for (int64_t t = 0; t < ntiles; ++t) {
        tot += tng[t];
        require(tot < (int64_t)INT32_MAX, IK_EINVAL, "mesh too large for int32 staging slots");
        tgptr[t + 1] = (int32_t)tot;
      }
      gq.resize(tot);
      parallelChunks(ntiles, [&](int64_t t0, int64_t, int64_t ch) {  // the same chunks as the loop above
        std::copy(gpart[ch].begin(), gpart[ch].end(), gq.begin() + tgptr[t0]);
      }, 256);
    }
    // An upper block that only one tile touches is complete after that tile's group sum, which the tile
    // writes straight into K (with the FULL mask and the mirror block). The others are staged, in gather
    // order: segments in the order qord, each ordered by tile. gdst[g] = slot, or -(K block + 1) with
    // gpt[g] = the mirror block and gmsk[g] = the mask of a direct group.
    HostVec<int32_t> ntile = filled<int32_t>(nup, 0);
    parallelChunks((int64_t)gq.size(), [&](int64_t g0, int64_t g1, int64_t) {
      for (int64_t g = g0; g < g1; ++g) __atomic_fetch_add(&ntile[gq[g]], 1, __ATOMIC_RELAXED);
    });
    parallelChunks(nup, [&](int64_t q0, int64_t q1, int64_t) {
      for (int64_t q = q0; q < q1; ++q) {
        gathered[q] = ntile[q] > 1;
        segCnt[q]   = gathered[q] ? ntile[q] : 0;
      }
    });
    h->nstage = layoutSegments();
    // A staged group's slot is its upper block's segment start plus its rank among that block's groups by
    // increasing group index (= tile order). Groups are dropped into their segment through a per-block cursor
    // (any order), each segment (at most 8 groups: one per tile touching the block) is then sorted, and the
    // k-th group of segment q gets slot segBeg[q] + k. A direct group stores -(K slot + 1), its mirror's K
    // slot and its FULL mask.
    const int64_t ngq = (int64_t)gq.size();
    HostVec<int32_t> gdst(ngq), gpt(ngq), cursor = filled<int32_t>(nup, 0), segG(h->nstage);
    HostVec<uint16_t> gmsk(ngq);
    parallelChunks(ngq, [&](int64_t g0, int64_t g1, int64_t) {
      for (int64_t g = g0; g < g1; ++g) {
        const int32_t q = gq[g];
        if (gathered[q]) {
          segG[segBeg[q] + __atomic_fetch_add(&cursor[q], 1, __ATOMIC_RELAXED)] = (int32_t)g;
          gpt[g]  = 0;
          gmsk[g] = 0;
        } else {
          gdst[g] = -(h->hpos[up[q]] + 1);
          gpt[g]  = h->hpos[upT[q]];
          gmsk[g] = mask[q];
        }
      }
    });
    parallelChunks(nup, [&](int64_t q0, int64_t q1, int64_t) {
      for (int64_t q = q0; q < q1; ++q) {
        if (!gathered[q]) continue;
        require(cursor[q] == segCnt[q], IK_EINVAL, "internal error: staging segment not filled");
        int32_t* b = segG.data() + segBeg[q];
        std::sort(b, b + segCnt[q]);
        for (int32_t k = 0; k < segCnt[q]; ++k) gdst[b[k]] = segBeg[q] + k;
      }
    });
    // the kernel's indexing relies on these: in every tile each valid entry has a distinct (group, rank)
    // with group < ng and rank < its group's size < tmaxr + 1, every group is non-empty, padding entries
    // target spare accumulators in [ng, 576); the slots are a permutation of [0, nstage) with each group
    // inside its own upper block's segment (the segments are disjoint consecutive ranges, each filled exactly
    // by its cursor), and every upper block is either staged by two or more tiles or written by exactly one
    {
      parallelChunks(nup, [&](int64_t q0, int64_t q1, int64_t) {
        for (int64_t q = q0; q < q1; ++q)
          require(ntile[q] >= 1 && segCnt[q] == (gathered[q] ? ntile[q] : 0) &&
                      cursor[q] == (gathered[q] ? segCnt[q] : 0),
                  IK_EINVAL, "internal error: upper block without a tile or a segment not filled");
      });
      parallelChunks((int64_t)gq.size(), [&](int64_t g0, int64_t g1, int64_t) {
        for (int64_t g = g0; g < g1; ++g) {
          const int32_t sl = gdst[g], q = gq[g];
          if (sl < 0)
            require(!gathered[q] && -sl - 1 == h->hpos[up[q]] && gpt[g] == h->hpos[upT[q]], IK_EINVAL,
                    "internal error: bad direct group");
          else
            require(sl >= segBeg[q] && sl < segBeg[q] + segCnt[q], IK_EINVAL,
                    "internal error: staging slot outside its segment");
        }
      });
      parallelChunks(ntiles, [&](int64_t t0, int64_t t1, int64_t) {
        std::vector<uint8_t> used(kTileEntries * 16);
        std::vector<int32_t> fill(kTileEntries);
        for (int64_t t = t0; t < t1; ++t) {
          const int ng = tgptr[t + 1] - tgptr[t];
          int nvalid   = 0;
          for (int s = 0; s < kTileElems; ++s) nvalid += tileElem[t * kTileElems + s] >= 0;
          require(tileElem[t * kTileElems] >= 0 && ng <= 36 * nvalid && tmaxr[t] >= 1 && tmaxr[t] <= 16,
                  IK_EINVAL, "internal error: bad tile");
          std::fill(used.begin(), used.begin() + ng * 16, 0);
          std::fill(fill.begin(), fill.begin() + ng, 0);
          const uint16_t* ep = epos.data() + t * kTileEntries;
          for (int j = 0; j < kTileEntries; ++j) {
            const int g = ep[j] & 0x3FF, rk = (ep[j] >> 10) & 0xF;
            if (j >= 36 * nvalid) {
              require(g >= ng && g < kTileEntries && rk == 0 && !(ep[j] & 0x8000u), IK_EINVAL,
                      "internal error: bad padding entry");
              continue;
            }
            require(g < ng && rk < tmaxr[t] && !used[g * 16 + rk], IK_EINVAL, "internal error: bad tile entry");
            used[g * 16 + rk] = 1;
            fill[g]++;
          }
          for (int g = 0; g < ng; ++g)
            for (int rk = 0; rk < 16; ++rk)
              require((rk < fill[g]) == (used[g * 16 + rk] != 0), IK_EINVAL, "internal error: gap in a tile group");
        }
      });
    }
    h->ntiles = ntiles;
    h->tileElem.upload(tileElem.data(), tileElem.size(), h->stream);
    h->tgptr.upload(tgptr.data(), tgptr.size(), h->stream);
    h->gdst.upload(gdst.data(), gdst.size(), h->stream);
    h->gpt.upload(gpt.data(), gpt.size(), h->stream);
    h->gmsk.upload(gmsk.data(), gmsk.size(), h->stream);
    h->epos.upload(epos.data(), epos.size(), h->stream);
    h->tmaxr.upload(tmaxr.data(), tmaxr.size(), h->stream);
  }
  // the gather's upper blocks: the gathered ones in gather order (chunkQ[c]: first one of chunk c), placed by
  // a prefix count over qord; each one's segment must start where the previous one's ends
  {
    HostVec<int64_t> at;
    parallelPrefix(nup, [&](int64_t k) { return (int64_t)gathered[qord[k]]; }, at);
    const int64_t ng = at[nup];
    HostVec<int32_t> cup(ng), cupT(ng), cseg(ng + 1);
    HostVec<uint16_t> cmask(ng);
    cseg[0] = 0;
    parallelChunks(nup, [&](int64_t k0, int64_t k1, int64_t) {
      for (int64_t k = k0; k < k1; ++k) {
        const int32_t q = qord[k];
        if (!gathered[q]) continue;
        const int64_t i = at[k];
        cup[i]          = h->hpos[up[q]];
        cupT[i]         = h->hpos[upT[q]];
        cmask[i]        = mask[q];
        cseg[i + 1]     = segBeg[q] + segCnt[q];
      }
    });
    parallelChunks(nup, [&](int64_t k0, int64_t k1, int64_t) {
      for (int64_t k = k0; k < k1; ++k)
        if (gathered[qord[k]])
          require(segBeg[qord[k]] == cseg[at[k]], IK_EINVAL, "internal error: gather segments not contiguous");
    });
    h->chunkQ.assign(nchunks + 1, 0);
    for (int64_t c = 0; c < nchunks; ++c) h->chunkQ[c + 1] = at[qchunk[c + 1]];
    h->upMask.upload(cmask.data(), cmask.size(), h->stream);
    h->upBlk.upload(cup.data(), cup.size(), h->stream);
    h->upT.upload(cupT.data(), cupT.size(), h->stream);
    h->segptr.upload(cseg.data(), cseg.size(), h->stream);
  }
  std::vector<int32_t> ncnt(nn + 1, 0);
  for (int64_t i = 0; i <= nn; ++i) ncnt[i] = (int32_t)cnt[i];
  h->rowptr.upload(h->hrowptr.data(), h->hrowptr.size(), h->stream);  // colidx: uploaded when ik_matrix_bsr asks
  {
    std::vector<int32_t> dslot(nn);
    for (int64_t i = 0; i < nn; ++i) dslot[i] = h->hpos[h->hdiag[i]];
    h->diagBlk.upload(dslot.data(), dslot.size(), h->stream);
  }
  // residual staging slot of every element corner: its position in the node's incidence list
  HostVec<int32_t> rdest((size_t)ne * 8);
  parallelChunks((int64_t)nlist.size(), [&](int64_t j0, int64_t j1, int64_t) {
    for (int64_t j = j0; j < j1; ++j) rdest[nlist[j]] = (int32_t)j;
  });
  h->ncptr.upload(ncnt.data(), ncnt.size(), h->stream);
  h->rdest.upload(rdest.data(), rdest.size(), h->stream);
  IK_HIP(hipStreamSynchronize(h->stream));
  h->patternBuilt = true;
}

}  // namespace ik
