// qd_internal.h -- shared state of libquadesth (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "quadesth.h"

namespace qd {

// Device bytes held by the library's buffers (all handles): current and
// high-water mark, for qd_memory_usage.
struct MemStats {
  std::atomic<long long> current{0}, peak{0};
  void add(long long b) {
    const long long c = current.fetch_add(b) + b;
    long long pk = peak.load();
    while (c > pk && !peak.compare_exchange_weak(pk, c)) {
    }
  }
};
inline MemStats& mem_stats() {
  static MemStats s;
  return s;
}

// Device buffer owned by the handle, or a view of memory it does not own
// (a slice of the allocation-free entry point's static workspace).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t count = 0;
  bool view = false;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p && !view) {
      (void)hipFree(p);
      mem_stats().add(-(long long)(count * sizeof(T)));
    }
    p = nullptr;
    count = 0;
    view = false;
  }
  // (Re)allocates when the size grows; contents are not preserved.  A view
  // never allocates: too small a view is an error.
  hipError_t ensure(size_t n) {
    if (n <= count && p) return hipSuccess;
    if (view) return hipErrorOutOfMemory;
    release();
    if (n == 0) return hipSuccess;
    hipError_t e = hipMalloc((void**)&p, n * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    count = n;
    mem_stats().add((long long)(n * sizeof(T)));
    return hipSuccess;
  }
};

// Per-element data the assembly kernels read with scalar (wave-uniform) loads.
struct ElemInfo {
  int node[9];   // 0-based node numbers
  int nconec;    // 9 or 6
  int kode;      // bit ld set when kode(e, ld) == 1 (traction prescribed)
  int k0mask;    // bit 3k+ld set when kode0(node k, ld) != 0 (assembly_mod.f90:149-161)
  int needG;     // G integrals contribute somewhere (a kode==0 dof or a non-zero bc)
  int pad[3];
};
static_assert(sizeof(ElemInfo) == 64, "ElemInfo layout");

// Singular (collocation node inside element) pair.
struct SingPair {
  int node;  // collocation node, 0-based
  int elem;  // element, 0-based
  int nodo;  // local node number of `node` in `elem`, 1-based
  int nsub;  // sub-triangles of the Duffy split (assembly_mod.f90:386-490)
};

// Sub-triangle table: per (element kind, nodo): up to 4 triangles of
// (v1[3], v2[3], area), assembly_mod.f90:399-489.
struct SubTri {
  double v1[3], v2[3], area, pad;
};

// Write plan of the assembly passes, kElemOps words per element: words 0-8
// are the H pass's column-block operations of the element's nodes, word 9
// the G pass's first-write bits (bit 3k+ld: column ld of node k is written
// for the first time, so it starts from zero instead of being read).
// An H-pass word: kOpFresh (the block's chain starts here: read A or start
// from zero), kOpLast (the chain ends here: store to A), else the block is
// carried to the chain's next element in LDS slot `out`; bits 2-4 the
// chain's written columns, bits 5-7 those it starts from zero, bits 8-11
// the slot it is carried in from, bits 12-15 the slot it is carried out in.
constexpr int kElemOps = 16;
constexpr int kCarrySlots = 4;
static_assert((kCarrySlots & (kCarrySlots - 1)) == 0, "carry slots: a power of two");
constexpr unsigned kOpFresh = 1u, kOpLast = 2u;
__host__ __device__ constexpr int op_wcols(unsigned op) { return (int)((op >> 2) & 7u); }
__host__ __device__ constexpr int op_zcols(unsigned op) { return (int)((op >> 5) & 7u); }
__host__ __device__ constexpr int op_slot_in(unsigned op) { return (int)((op >> 8) & 15u); }
__host__ __device__ constexpr int op_slot_out(unsigned op) { return (int)((op >> 12) & 15u); }

constexpr int kGP = 36;         // Gauss points per element slot (6x6; tri6 uses 16)
constexpr int kAsmStride = 16;  // doubles per Gauss point in the assembly table
constexpr int kAsmStride32 = 16;  // floats per Gauss point in the FP32 far-field table
constexpr int kNearGL = 40;     // near-field rule of the internal points (stress_mod.f90:513)

struct Material {
  double mu, nu, young, c3, c4, c5, c6;
};

}  // namespace qd

struct qd_lu_state;

struct qd_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // set for the duration of an allocation-free call: QD_ALLOC then fails
  // instead of growing a buffer
  bool forbid_alloc = false;

  // problem data (host copies)
  int n = 0, ne = 0;
  std::vector<double> hx, hy, hz;
  std::vector<int> hnconec, hkconec, hkode;
  std::vector<double> hbc;
  std::vector<int> nesim;      // elements sharing each node
  std::vector<int> ke_ptr, ke;  // elements of node i: ke[ke_ptr[i] .. ke_ptr[i+1]), 1-based, ascending
  qd::Material mat{};
  bool have_mesh = false, have_mat = false, have_bc = false, have_geom = false, have_sys = false;

  // device mesh
  qd::DevBuf<double> x, y, z;
  qd::DevBuf<int> nconec, kconec0;  // kconec0: 0-based, (ne, 9)
  qd::DevBuf<double> geo;           // reference tables: 8 blocks of (36, ne) + ff (9, 36, ne)
  qd::DevBuf<double> asmtab;        // (ne*36, 16): xg(3) eta(3) p12(9) pad
  qd::DevBuf<float> asmtab32;       // (ne*36, 16): eta(3) p12(9) pad, FP32
  qd::DevBuf<double> elfar;         // (ne, 4): centroid, squared far-field distance
  qd::DevBuf<qd::ElemInfo> einfo;
  qd::DevBuf<unsigned> elop;        // (ne, kElemOps): the assembly's write plan
  std::vector<int> h_unwritten;     // columns of A that no pass writes (zero-filled)
  qd::DevBuf<int> unwritten;
  qd::DevBuf<double> ebc, evp;      // (ne, 27): bc and vp per element dof
  qd::DevBuf<int> node_k0;          // (n, 3) kode0 flag per node/direction
  qd::DevBuf<double> node_vp;       // (n, 3)
  qd::DevBuf<qd::SubTri> subtri;    // [2 kinds][10 nodo][4]

  // assembly schedule
  int chunk_elems = 32;
  std::vector<int> h_chunk_start;                 // nchunk + 1
  std::vector<int> h_colour_chunks;               // chunk ids sorted by colour
  std::vector<int> h_colour_start;                // ncolour + 1 offsets into colour_chunks
  std::vector<int> h_gchunks;                     // chunks holding needG elements, colour order
  std::vector<int> h_gcolour_start;
  qd::DevBuf<int> chunk_start, colour_chunks, gchunks;
  std::vector<qd::SingPair> h_pairs;              // sorted by node, then element
  std::vector<int> h_pair_ptr;                    // n + 1
  qd::DevBuf<qd::SingPair> pairs;
  qd::DevBuf<int> pair_ptr;
  qd::DevBuf<double> sing_hg;                     // (npairs, 162)
  qd::DevBuf<double> dpart, fpart_h, fpart_g;     // assembly partial sums

  // assembled system (column-major, leading dimension lda)
  qd::DevBuf<double> A, FI;
  long long lda = 0;
  // QD_ASSEMBLY_*: far-field pairs in FP32 for A itself (1) or only for the
  // matrix the mixed-precision factorisations factor (2, in A_pre)
  int asm_mode = QD_ASSEMBLY_FP64;
  qd::DevBuf<double> A_pre, FI_pre;
  bool have_pre = false;

  // solver state (qd_lu.hip)
  qd_lu_state* lu = nullptr;

  // Somigliana tables
  qd::DevBuf<double> far_tab, near_tab, el_centre, somi_part;
  qd::DevBuf<int> far_start;
  qd::DevBuf<double> tmp_fi, tmp_dfi, tmp_pts, tmp_out;
  qd::DevBuf<long long> counters;
  // near point-element pairs: point-major list (ptr, element, point), the
  // element-major index, the near pass's work items and per-pair sums
  qd::DevBuf<int> near_cnt, near_ptr, near_pe, near_pp, near_ecnt, near_efill, near_eptr, near_ep;
  qd::DevBuf<int> near_cch, near_cptr;
  qd::DevBuf<int2> near_items;
  qd::DevBuf<double> near_out;
  qd::DevBuf<long long> near_total;

  // instrumentation (qd_profile_enable / qd_profile_read)
  struct ProfRec {
    int kind;
    hipEvent_t a, b;
    double work;
  };
  unsigned prof_mask = 0;
  std::vector<ProfRec> prof_pending;
  long long prof_launches[QD_PROF_KINDS] = {};
  double prof_ms[QD_PROF_KINDS] = {};
  double prof_work[QD_PROF_KINDS] = {};
};

namespace qd {
// Brackets a launch of class `kind` with events when profiling is on.
struct ProfScope {
  qd_context* h;
  int kind;
  double work;
  hipStream_t st;
  hipEvent_t a = nullptr;
  ProfScope(qd_context* h_, int kind_, double work_) : ProfScope(h_, kind_, work_, h_->stream) {}
  ProfScope(qd_context* h_, int kind_, double work_, hipStream_t st_) : h(h_), kind(kind_), work(work_), st(st_) {
    if (kind >= 0 && ((h->prof_mask >> kind) & 1)) {
      if (hipEventCreate(&a) == hipSuccess) (void)hipEventRecord(a, st);
      else a = nullptr;
    }
  }
  ~ProfScope() {
    if (!a) return;
    hipEvent_t b;
    if (hipEventCreate(&b) != hipSuccess) return;
    (void)hipEventRecord(b, st);
    h->prof_pending.push_back(qd_context::ProfRec{kind, a, b, work});
  }
};
}  // namespace qd

namespace qd {
int set_error(qd_context* h, int code, const char* fmt, ...);
#define QD_HIP(h, call)                                                                 \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return qd::set_error((h), QD_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call, \
                           hipGetErrorString(e_));                                      \
  } while (0)
#define QD_ALLOC(h, buf, n)                                                                    \
  do {                                                                                         \
    if ((h)->forbid_alloc && ((buf).count < (size_t)(n) || !(buf).p))                           \
      return qd::set_error((h), QD_ERR_NOMEM, "%s:%d: %zu elements needed in an allocation-free call", \
                           __FILE__, __LINE__, (size_t)(n));                                    \
    hipError_t e_ = (buf).ensure(n);                                                           \
    if (e_ != hipSuccess)                                                                      \
      return qd::set_error((h), QD_ERR_NOMEM, "%s:%d allocating %zu elements: %s", __FILE__, \
                           __LINE__, (size_t)(n), hipGetErrorString(e_));                     \
  } while (0)

// phases implemented in the .hip files
int geometry_run(qd_context* h);
// host-side assembly preprocessing (qd_plan.cpp)
void build_schedule(qd_context* h);
int bc_tables(qd_context* h, const int* kode, const double* bc, std::vector<ElemInfo>& info,
              std::vector<double>& evp, std::vector<int>& nk0, std::vector<double>& nvp);
void build_gschedule(qd_context* h, const std::vector<ElemInfo>& info);
void build_write_plan(qd_context* h, const std::vector<ElemInfo>& info, std::vector<unsigned>& ops);
// A (or, with into_pre, A_pre) and FI (or fi_out_dev); far32: far-field pairs in FP32.
int assembly_run(qd_context* h, double* fi_out_dev, bool write_A, bool far32 = false, bool into_pre = false);
int somigliana_run(qd_context* h, int l, const double* pts_dev, const double* fi_dev,
                   const double* dfi_dev, int variant, double* u_dev, double* sig_dev,
                   double* prin_dev, double* vm_dev, long long* counts);
// Element centroids and near thresholds (stress_mod.f90:584-604) into h->el_centre.
int somigliana_centres(qd_context* h);
// 1-D Gauss-Legendre rule of stress_mod.f90:475-503 (host).
void gauss_legendre(int npts, double* xg, double* wg);
// Reverse-mode derivatives (qd_grad.hip).
int ghmat8_vjp_run(qd_context* h, const double* lambda, const double* fi, const double* dfi, double* gx,
                   double* gy, double* gz, double* gnu);
int somigliana_vjp_run(qd_context* h, int l, const double* pts, const double* fi, const double* dfi,
                       int variant, const double* gu, const double* gs, double* gfi, double* gdfi, double* gx,
                       double* gy, double* gz, double* gnu);
int boundary_stress_run(qd_context* h, const double* fi_dev, const double* dfi_dev,
                        double* se_dev, double* sn_dev, double* prin_dev, double* vm_dev);
int lu_solve(qd_context* h, int precision, int nrhs, double* b_dev, long long ldb, double tol,
             int max_iter, bool transpose, qd_solve_report* rep);
// lu_solve with the held factors and every work vector in the library's
// static device workspace (no allocation); QD_ERR_STATE without such factors.
int lu_solve_noalloc(qd_context* h, int precision, int nrhs, double* b_dev, long long ldb, double tol,
                     int max_iter, qd_solve_report* rep);
// Once per device, with that device current: the workspace's device address.
int noalloc_init(int device, std::string* err);
void lu_free(qd_context* h);
int fp16_update(qd_context* h, long long M, long long N, int Kp, const void* A16, const void* B16,
                const float* colscale, float* C, long long ldc);
int lu_invalidate(qd_context* h);
int dgesv_device(qd_context* h, int n, int nrhs, double* a_dev, long long lda, int* ipiv_host,
                 double* b_dev, long long ldb);
}  // namespace qd
