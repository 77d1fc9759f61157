// qd_plan.cpp -- host-side preprocessing of the assembly: the chunk
// schedule and its colouring, the boundary-condition tables of the element
// passes, and the write plan by which each column block of A is written once
// per chain of element updates.  Host code only (no device calls), so a CPU
// test can drive it.
#include <algorithm>
#include <vector>

#include "qd_internal.h"

namespace qd {

// Greedy colouring of element chunks so that no two chunks of one colour
// share a node: chunks of one colour can then update the same rows of A
// without write conflicts.
void build_schedule(qd_context* h) {
  const int ne = h->ne, n = h->n;
  int P = std::max(8, std::min(64, ne / 256));
  h->chunk_elems = P;
  h->h_chunk_start.clear();
  int e = 0;
  while (e < ne) {
    h->h_chunk_start.push_back(e);
    int first = e;
    int cnt = 0;
    while (e < ne && cnt < P && h->hnconec[e] == h->hnconec[first]) {
      ++e;
      ++cnt;
    }
  }
  const int nchunk = (int)h->h_chunk_start.size();
  h->h_chunk_start.push_back(ne);
  std::vector<std::vector<int>> node_chunks(n);
  for (int c = 0; c < nchunk; ++c)
    for (int el = h->h_chunk_start[c]; el < h->h_chunk_start[c + 1]; ++el)
      for (int k = 0; k < h->hnconec[el]; ++k) {
        int nd = h->hkconec[el * 9 + k] - 1;
        auto& v = node_chunks[nd];
        if (v.empty() || v.back() != c) v.push_back(c);
      }
  std::vector<int> colour(nchunk, -1);
  int ncol = 0;
  std::vector<char> used;
  for (int c = 0; c < nchunk; ++c) {
    used.assign(ncol + 1, 0);
    for (int el = h->h_chunk_start[c]; el < h->h_chunk_start[c + 1]; ++el)
      for (int k = 0; k < h->hnconec[el]; ++k)
        for (int oc : node_chunks[h->hkconec[el * 9 + k] - 1])
          if (colour[oc] >= 0) used[colour[oc]] = 1;
    int col = 0;
    while (col < ncol && used[col]) ++col;
    colour[c] = col;
    if (col == ncol) ++ncol;
  }
  h->h_colour_start.assign(ncol + 1, 0);
  h->h_colour_chunks.clear();
  for (int col = 0; col < ncol; ++col) {
    h->h_colour_start[col] = (int)h->h_colour_chunks.size();
    for (int c = 0; c < nchunk; ++c)
      if (colour[c] == col) h->h_colour_chunks.push_back(c);
  }
  h->h_colour_start[ncol] = (int)h->h_colour_chunks.size();
}

// kode0 / vp search of assembly_mod.f90:149-161 for node nd (0-based).
static int find_vp(qd_context* h, const int* kode, const double* bc, int nd, int ld, double* vp) {
  for (int i = 0; i < h->nesim[nd]; ++i) {
    int e = h->ke[h->ke_ptr[nd] + i] - 1;
    if (kode[e * 3 + ld] == 0) {
      for (int m = 0; m < h->hnconec[e]; ++m)
        if (h->hkconec[e * 9 + m] - 1 == nd) {
          *vp = bc[e * 27 + 3 * m + ld];
          return 1;
        }
    }
  }
  *vp = 0.0;
  return 0;
}

int bc_tables(qd_context* h, const int* kode, const double* bc, std::vector<ElemInfo>& info,
                     std::vector<double>& evp, std::vector<int>& nk0, std::vector<double>& nvp) {
  const int n = h->n, ne = h->ne;
  nk0.assign(3 * (size_t)n, 0);
  nvp.assign(3 * (size_t)n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int ld = 0; ld < 3; ++ld) nk0[3 * i + ld] = find_vp(h, kode, bc, i, ld, &nvp[3 * i + ld]);
  info.assign(ne, ElemInfo{});
  evp.assign(27 * (size_t)ne, 0.0);
  for (int e = 0; e < ne; ++e) {
    ElemInfo& f = info[e];
    f.nconec = h->hnconec[e];
    f.kode = 0;
    f.k0mask = 0;
    f.needG = 0;
    for (int ld = 0; ld < 3; ++ld) {
      if (kode[e * 3 + ld] != 0 && kode[e * 3 + ld] != 1)
        return set_error(h, QD_ERR_ARG, "kode(%d) = %d (expected 0 or 1)", 3 * e + ld + 1,
                         kode[e * 3 + ld]);
      if (kode[e * 3 + ld] == 1) f.kode |= 1 << ld;
      else f.needG = 1;
    }
    for (int k = 0; k < 9; ++k) f.node[k] = (k < f.nconec) ? h->hkconec[e * 9 + k] - 1 : -1;
    for (int k = 0; k < f.nconec; ++k)
      for (int ld = 0; ld < 3; ++ld) {
        int nd = f.node[k];
        if (nk0[3 * nd + ld]) f.k0mask |= 1 << (3 * k + ld);
        evp[e * 27 + 3 * k + ld] = nvp[3 * nd + ld];
        if (bc[e * 27 + 3 * k + ld] != 0.0 && (f.kode >> ld & 1)) f.needG = 1;
      }
  }
  return QD_OK;
}

// G-pass schedule: the chunks holding needG elements, in colour order.
void build_gschedule(qd_context* h, const std::vector<ElemInfo>& info) {
  h->h_gchunks.clear();
  h->h_gcolour_start.assign(h->h_colour_start.size(), 0);
  for (size_t col = 0; col + 1 < h->h_colour_start.size(); ++col) {
    h->h_gcolour_start[col] = (int)h->h_gchunks.size();
    for (int i = h->h_colour_start[col]; i < h->h_colour_start[col + 1]; ++i) {
      int c = h->h_colour_chunks[i];
      bool any = false;
      for (int e = h->h_chunk_start[c]; e < h->h_chunk_start[c + 1]; ++e) any |= info[e].needG != 0;
      if (any) h->h_gchunks.push_back(c);
    }
  }
  h->h_gcolour_start.back() = (int)h->h_gchunks.size();
}

// Write plan of the assembly passes (the kElemOps words per element).  The
// launch order is colour by colour, a colour's H runs before its G runs, and
// a chunk's elements in order; chunks of one launch share no node.  The H
// updates of one node's column block inside a chunk form chains: a chain
// starts from zero in the columns no earlier update (in launch order) wrote
// and from A in the others, is carried between its elements in an LDS slot,
// and is stored once at its end.  Which gaps between a node's consecutive
// elements are carried is an interval schedule over kCarrySlots slots:
// gaps by earliest end, each into the free slot that was released last,
// which carries the largest number of gaps; a slot read at an element is
// free again from the next element on.  A G update starts from zero where
// it is its column's first write.  Columns that no update writes are listed
// in h_unwritten (zero-filled before the passes).
void build_write_plan(qd_context* h, const std::vector<ElemInfo>& info, std::vector<unsigned>& ops) {
  const int n = h->n, ne = h->ne;
  ops.assign((size_t)ne * kElemOps, 0u);
  std::vector<unsigned char> written(n, 0);
  auto hcols = [&](int e, int k) {  // columns the H pass writes for node k of e
    int m = 0;
    for (int ld = 0; ld < 3; ++ld)
      if (((info[e].kode >> ld) & 1) && !((info[e].k0mask >> (3 * k + ld)) & 1)) m |= 1 << ld;
    return m;
  };
  struct App {
    int node, t, k;
  };
  std::vector<App> app;
  std::vector<int> gap_slot;
  struct Gap {
    int start, end, idx;
  };
  std::vector<Gap> gaps;
  const int ncol = (int)h->h_colour_start.size() - 1;
  for (int col = 0; col < ncol; ++col) {
    for (int i = h->h_colour_start[col]; i < h->h_colour_start[col + 1]; ++i) {
      const int c = h->h_colour_chunks[i];
      const int e0 = h->h_chunk_start[c], e1 = h->h_chunk_start[c + 1];
      app.clear();
      for (int e = e0; e < e1; ++e)
        for (int k = 0; k < info[e].nconec; ++k) app.push_back(App{info[e].node[k], e - e0, k});
      std::stable_sort(app.begin(), app.end(), [](const App& a, const App& b) { return a.node < b.node; });
      // gap j joins app[j] and app[j + 1] of the same node
      gaps.clear();
      gap_slot.assign(app.size(), -1);
      for (size_t j = 0; j + 1 < app.size(); ++j)
        if (app[j + 1].node == app[j].node) gaps.push_back(Gap{app[j].t, app[j + 1].t, (int)j});
      std::stable_sort(gaps.begin(), gaps.end(), [](const Gap& a, const Gap& b) { return a.end < b.end; });
      int slot_end[kCarrySlots];
      for (int s = 0; s < kCarrySlots; ++s) slot_end[s] = -1;
      for (const Gap& g : gaps) {
        int best = -1;
        for (int s = 0; s < kCarrySlots; ++s)
          if (slot_end[s] < g.start && (best < 0 || slot_end[s] > slot_end[best])) best = s;
        if (best >= 0) {
          slot_end[best] = g.end;
          gap_slot[g.idx] = best;
        }
      }
      // chains, per node in element order
      size_t j = 0;
      while (j < app.size()) {
        size_t b = j;
        int wc = 0;
        while (true) {
          wc |= hcols(e0 + app[j].t, app[j].k);
          if (gap_slot[j] < 0) break;
          ++j;
        }
        const int nd = app[b].node;
        const int z = wc & ~written[nd];
        written[nd] |= (unsigned char)wc;
        for (size_t q = b; q <= j; ++q) {
          unsigned op = ((unsigned)wc << 2) | ((unsigned)z << 5);
          if (q == b) op |= kOpFresh;
          else op |= (unsigned)gap_slot[q - 1] << 8;
          if (q == j) op |= kOpLast;
          else op |= (unsigned)gap_slot[q] << 12;
          ops[(size_t)(e0 + app[q].t) * kElemOps + app[q].k] = op;
        }
        ++j;
      }
    }
    for (int i = h->h_gcolour_start[col]; i < h->h_gcolour_start[col + 1]; ++i) {
      const int c = h->h_gchunks[i];
      for (int e = h->h_chunk_start[c]; e < h->h_chunk_start[c + 1]; ++e) {
        if (!info[e].needG) continue;
        unsigned gz = 0;
        for (int k = 0; k < info[e].nconec; ++k)
          for (int ld = 0; ld < 3; ++ld) {
            if ((info[e].kode >> ld) & 1) continue;
            const int nd = info[e].node[k];
            if (!((written[nd] >> ld) & 1)) gz |= 1u << (3 * k + ld);
            written[nd] |= (unsigned char)(1 << ld);
          }
        ops[(size_t)e * kElemOps + 9] = gz;
      }
    }
  }
  h->h_unwritten.clear();
  for (int nd = 0; nd < n; ++nd)
    for (int ld = 0; ld < 3; ++ld)
      if (!((written[nd] >> ld) & 1)) h->h_unwritten.push_back(3 * nd + ld);
}

}  // namespace qd
