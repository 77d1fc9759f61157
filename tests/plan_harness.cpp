// Host-only driver of bem3d_elastostatic_amd/csrc/qd_plan.cpp for
// tests/test_write_plan.py.  Reads a mesh and its boundary conditions from
// the binary files in the directory given, builds the chunk schedule, the
// boundary-condition tables and the write plan as qd_set_mesh / qd_set_bc
// do, then replays the assembly passes on the CPU for a sample of
// collocation rows twice, with a distinct pseudo-random value per element
// contribution:
//  * as plain read-modify-write updates of a zeroed matrix, in launch order
//    (colour by colour, H chunks then G chunks, elements in chunk order);
//  * by the write plan: A starts as NaN, the unwritten columns are zeroed,
//    H chains start from zero or from A, travel in per-lane carry slots whose
//    owner is checked at every read, and are stored at their ends; G updates
//    start from zero where the plan says so.
// The two matrices must agree bit for bit in every sampled row.  Also checks
// that no two chunks of one colour share a node.  No device is used.
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "qd_internal.h"

namespace qd {
int set_error(qd_context*, int code, const char* fmt, ...) {
  char b[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(b, sizeof b, fmt, ap);
  va_end(ap);
  std::fprintf(stderr, "error %d: %s\n", code, b);
  return code;
}
}  // namespace qd

using namespace qd;

template <class T>
static std::vector<T> rd(const std::string& p) {
  std::ifstream f(p, std::ios::binary | std::ios::ate);
  const size_t n = (size_t)f.tellg();
  f.seekg(0);
  std::vector<T> v(n / sizeof(T));
  f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)n);
  return v;
}

// a distinct value in [-1, 1) per (pass, row node, element, node slot, ld, q)
static double val(uint64_t pass, uint64_t in, uint64_t e, uint64_t k, uint64_t ld, uint64_t q) {
  uint64_t x = pass * 0x9E3779B97F4A7C15ull ^ (in + 1) * 0xBF58476D1CE4E5B9ull ^ (e + 7) * 0x94D049BB133111EBull ^
               (k * 9 + ld * 3 + q + 1) * 0xD6E8FEB86659FD93ull;
  x ^= x >> 31;
  x *= 0x9E3779B97F4A7C15ull;
  x ^= x >> 29;
  return (double)(x >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string d = argv[1];
  qd_context h;
  std::ifstream m(d + "/meta.txt");
  int n, ne;
  m >> n >> ne;
  h.n = n;
  h.ne = ne;
  h.hnconec = rd<int>(d + "/nconec.bin");
  h.hkconec = rd<int>(d + "/kconec.bin");
  const std::vector<int> kode = rd<int>(d + "/kode.bin");
  const std::vector<double> bc = rd<double>(d + "/bc.bin");
  const std::vector<int> rows = rd<int>(d + "/rows.bin");
  // adjacency as qd_set_mesh builds it
  h.nesim.assign(n, 0);
  for (int e = 0; e < ne; ++e)
    for (int j = 0; j < h.hnconec[e]; ++j) h.nesim[h.hkconec[e * 9 + j] - 1] += 1;
  h.ke_ptr.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) h.ke_ptr[i + 1] = h.ke_ptr[i] + h.nesim[i];
  h.ke.assign((size_t)h.ke_ptr[n], 0);
  std::vector<int> fill(h.ke_ptr.begin(), h.ke_ptr.end() - 1);
  for (int e = 0; e < ne; ++e)
    for (int j = 0; j < h.hnconec[e]; ++j) h.ke[fill[h.hkconec[e * 9 + j] - 1]++] = e + 1;
  build_schedule(&h);
  std::vector<ElemInfo> info;
  std::vector<double> evp, nvp;
  std::vector<int> nk0;
  if (bc_tables(&h, kode.data(), bc.data(), info, evp, nk0, nvp)) return 1;
  build_gschedule(&h, info);
  std::vector<unsigned> ops;
  build_write_plan(&h, info, ops);

  const int ncol = (int)h.h_colour_start.size() - 1;
  // colouring: chunks of one colour share no node
  for (int col = 0; col < ncol; ++col) {
    std::vector<int> owner(n, -1);
    for (int i = h.h_colour_start[col]; i < h.h_colour_start[col + 1]; ++i) {
      const int c = h.h_colour_chunks[i];
      for (int e = h.h_chunk_start[c]; e < h.h_chunk_start[c + 1]; ++e)
        for (int k = 0; k < info[e].nconec; ++k) {
          int& o = owner[info[e].node[k]];
          if (o >= 0 && o != c) {
            std::printf("colour %d: chunks %d and %d share node %d\n", col, o, c, info[e].node[k]);
            return 1;
          }
          o = c;
        }
    }
  }

  const long long N = 3LL * n;
  const int R = (int)rows.size();
  const double mu = 0.37;
  // A restricted to the sampled rows: column c, sampled row r, component q
  auto at = [&](std::vector<double>& A, long long c, int r, int q) -> double& { return A[(c * R + r) * 3 + q]; };
  std::vector<double> Aref((size_t)N * R * 3, 0.0), Aplan((size_t)N * R * 3, std::nan(""));
  for (int c : h.h_unwritten)
    for (int r = 0; r < R; ++r)
      for (int q = 0; q < 3; ++q) at(Aplan, c, r, q) = 0.0;
  auto is_sing = [&](int e, int in) {
    for (int k = 0; k < info[e].nconec; ++k)
      if (info[e].node[k] == in) return true;
    return false;
  };
  auto hwrites = [&](int e, int k, int ld) {
    return ((info[e].kode >> ld) & 1) && !((info[e].k0mask >> (3 * k + ld)) & 1);
  };
  long long chains = 0, carried = 0, loads = 0;
  std::vector<double> carry((size_t)R * kCarrySlots * 9);
  int slot_owner[kCarrySlots];
  for (int col = 0; col < ncol; ++col) {
    for (int i = h.h_colour_start[col]; i < h.h_colour_start[col + 1]; ++i) {
      const int c = h.h_colour_chunks[i];
      for (int s = 0; s < kCarrySlots; ++s) slot_owner[s] = -1;
      std::fill(carry.begin(), carry.end(), std::nan(""));
      for (int e = h.h_chunk_start[c]; e < h.h_chunk_start[c + 1]; ++e) {
        const ElemInfo& f = info[e];
        for (int k = 0; k < f.nconec; ++k) {
          const int ik = f.node[k];
          const unsigned op = ops[(size_t)e * kElemOps + k];
          const int wc = op_wcols(op), z = op_zcols(op);
          int want = 0;
          for (int ld = 0; ld < 3; ++ld) want |= hwrites(e, k, ld) << ld;
          if ((want & ~wc) || (z & ~wc)) {
            std::printf("element %d node %d: plan columns %d/%d, element writes %d\n", e, k, wc, z, want);
            return 1;
          }
          if (op & kOpFresh) {
            ++chains;
            for (int ld = 0; ld < 3; ++ld) loads += ((wc & ~z) >> ld) & 1;
          } else {
            const int s = op_slot_in(op);
            if (s >= kCarrySlots || slot_owner[s] != ik) {
              std::printf("element %d node %d: carry slot %d held by node %d\n", e, ik, s,
                          s < kCarrySlots ? slot_owner[s] : -2);
              return 1;
            }
            slot_owner[s] = -1;
            ++carried;
          }
          int so = -1;
          if (!(op & kOpLast)) {
            so = op_slot_out(op);
            if (so >= kCarrySlots || slot_owner[so] != -1) {
              std::printf("element %d node %d: carry slot %d busy\n", e, ik, so);
              return 1;
            }
            slot_owner[so] = ik;
          }
          for (int r = 0; r < R; ++r) {
            const int in = rows[r];
            const bool sing = is_sing(e, in);
            double v[3][3];
            for (int ld = 0; ld < 3; ++ld)
              for (int q = 0; q < 3; ++q) {
                if (op & kOpFresh)
                  v[ld][q] = (((wc & ~z) >> ld) & 1) ? at(Aplan, 3LL * ik + ld, r, q) : 0.0;
                else
                  v[ld][q] = carry[((size_t)r * kCarrySlots + op_slot_in(op)) * 9 + 3 * ld + q];
                if (!sing && hwrites(e, k, ld)) {
                  const double hv = val(0, in, e, k, ld, q);
                  v[ld][q] += hv;
                  at(Aref, 3LL * ik + ld, r, q) += hv;
                }
              }
            for (int ld = 0; ld < 3; ++ld)
              for (int q = 0; q < 3; ++q) {
                if (op & kOpLast) {
                  if ((wc >> ld) & 1) at(Aplan, 3LL * ik + ld, r, q) = v[ld][q];
                } else {
                  carry[((size_t)r * kCarrySlots + so) * 9 + 3 * ld + q] = v[ld][q];
                }
              }
          }
        }
      }
      for (int s = 0; s < kCarrySlots; ++s)
        if (slot_owner[s] != -1) {
          std::printf("chunk %d ends with slot %d held by node %d\n", c, s, slot_owner[s]);
          return 1;
        }
    }
    for (int i = h.h_gcolour_start[col]; i < h.h_gcolour_start[col + 1]; ++i) {
      const int c = h.h_gchunks[i];
      for (int e = h.h_chunk_start[c]; e < h.h_chunk_start[c + 1]; ++e) {
        const ElemInfo& f = info[e];
        if (!f.needG) continue;
        const unsigned gz = ops[(size_t)e * kElemOps + 9];
        for (int k = 0; k < f.nconec; ++k)
          for (int ld = 0; ld < 3; ++ld) {
            if ((f.kode >> ld) & 1) continue;
            const long long cc = 3LL * f.node[k] + ld;
            const bool z = (gz >> (3 * k + ld)) & 1;
            for (int r = 0; r < R; ++r) {
              const int in = rows[r];
              if (is_sing(e, in)) {
                if (z)
                  for (int q = 0; q < 3; ++q) at(Aplan, cc, r, q) = 0.0;
                continue;
              }
              for (int q = 0; q < 3; ++q) {
                const double g = val(1, in, e, k, ld, q);
                at(Aref, cc, r, q) = at(Aref, cc, r, q) - g * mu;
                at(Aplan, cc, r, q) = (z ? 0.0 : at(Aplan, cc, r, q)) - g * mu;
              }
            }
          }
      }
    }
  }
  long long bad = 0;
  for (long long cc = 0; cc < N; ++cc)
    for (int r = 0; r < R; ++r)
      for (int q = 0; q < 3; ++q)
        if (std::memcmp(&at(Aref, cc, r, q), &at(Aplan, cc, r, q), sizeof(double)) != 0) {
          if (bad < 5)
            std::printf("column %lld row %d.%d: sequential %.17g, plan %.17g\n", cc, rows[r], q, at(Aref, cc, r, q),
                        at(Aplan, cc, r, q));
          ++bad;
        }
  long long incid = 0;
  for (int e = 0; e < ne; ++e) incid += info[e].nconec;
  std::printf("incidences %lld chains %lld carried %lld column_loads %lld unwritten %zu mismatches %lld\n", incid,
              chains, carried, loads, h.h_unwritten.size(), bad);
  return bad ? 1 : 0;
}
