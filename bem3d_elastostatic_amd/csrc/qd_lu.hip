// qd_lu.hip -- dense LU with partial pivoting, triangular solves and mixed
// precision iterative refinement on gfx950 (the DGESV of solver_mod.f90:28).
//
// Layout.  The assembled A is column-major (the reference's layout, written
// coalesced by the assembly).  The factorisation works on a row-major copy W
// (one transpose on the way in): row interchanges of partial pivoting then
// move contiguous rows, and every GEMM/TRSM tile load is a set of contiguous
// row segments.
//
// Algorithm.  Right-looking blocked LU, panel width NB:
//   panel   W[j:N, j:j+NB] factored recursively (halves: left LU, TRSM,
//           GEMM update, right LU); leaves of width <= 32 advance one column
//           per launch: the launch scales the previous column's multipliers,
//           applies its rank-1 update to the leaf, finds the next pivot (block
//           maxima, the last block to finish reduces them, the first largest
//           |a| wins as in LAPACK idamax) and swaps the panel rows;
//   laswp   the panel's interchanges applied to the other columns;
//   trsm    U12 = L11^-1 A12;
//   gemm    A22 -= L21 U12 with MFMA (f64 16x16x4, f32 16x16x4, or
//           f16 16x16x32 with FP32 accumulation and per-64-column scaling of
//           U12; L21 needs none, |l| <= 1 under partial pivoting).
// The factors are LAPACK's: P A = L U with ipiv the sequence of interchanges.
//
// Refinement.  x <- x + LU^-1 (b - A x): the residual always uses the FP64
// column-major A; the correction uses the low-precision factors.  When the
// plain iteration stalls, restarted GMRES preconditioned by the same factors
// (FP64 Arnoldi) takes over.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

#include "qd_internal.h"

struct qd_lu_state {
  int precision = -1;  // precision of the factors held
  bool valid = false;
  int info = 0;
  long long N = 0, ld = 0;
  qd::DevBuf<double> W64;
  qd::DevBuf<float> W32;
  qd::DevBuf<int> ipiv_dev;  // 0-based absolute pivot rows
  qd::DevBuf<int> perm_dev;
  std::vector<int> ipiv;
  qd::DevBuf<double> piv_part;   // block partial maxima
  qd::DevBuf<int> piv_idx;
  qd::DevBuf<int> piv_state;     // [0] counter, [1] info
  qd::DevBuf<double> rpiv;       // reciprocal of the current pivot
  // FP16 operand copies of the GEMMs (fp16 mode): one set for the outer
  // step's updates, one for the panel, which runs on its own stream.
  struct F16Scratch {
    qd::DevBuf<float> colscale;  // per-64-column scales of U12
    qd::DevBuf<_Float16> h16a, h16b;
  };
  F16Scratch sc_main, sc_panel;
  F16Scratch sc_solve;  // the inner updates of the main stream's U12 solve, so sc_main keeps L21
  F16Scratch sc_u12;    // those of the aux stream's U12 solve, which can overlap the main one
  F16Scratch sc_l21;    // the product L21 = A21 U11^-1 of a diagonal-block panel (aux stream)
  qd::DevBuf<float> binv;   // a diagonal-block panel's U11^-1 (npad x npad)
  qd::DevBuf<float> binvl;  // and its L11^-1
  qd::DevBuf<float> binv_t; // their recursive doubling's intermediate products
  qd::DevBuf<float> bdiag;  // u_kk, then -1 / u_kk (2 x npad)
  qd::DevBuf<float> bstep;  // step buffers of the diagonal block's 64-column steps
  // FP16 copies of the L21 that the last two diagonal-block panels wrote (the
  // trailing updates' left operand), and the first column of their panels
  qd::DevBuf<_Float16> l21h[2];
  long long l21h_col[2] = {-1, -1};
  int l21h_next = 0;
  // likewise the strictly lower part of L11^-1, negated, in FP16 ([jb][Kp]),
  // for the U12 solves of those panels' steps, with its panel's first column
  // and width
  qd::DevBuf<_Float16> linvh[2];
  long long linvh_col[2] = {-1, -1};
  int linvh_w[2] = {0, 0};
  hipStream_t aux = nullptr;     // panel stream of the look-ahead (non-blocking, default priority)
  hipEvent_t ev_cols = nullptr, ev_panel = nullptr;
  hipEvent_t ev_top = nullptr, ev_u12 = nullptr;  // split steps: next U12 rows updated / solved
  qd::DevBuf<double> vec;        // solve / refinement work vectors
  qd::DevBuf<float> vecf;
  qd::DevBuf<double> gm;         // GMRES basis
  qd::DevBuf<double> rowpart;    // residual slice partials
  qd::DevBuf<double> nrm;        // block maxima of the inf-norms
  qd::DevBuf<int> ts_cand;       // tournament candidate lists, all levels
  qd::DevBuf<unsigned> ts_cnt;   // tournament group counters (zero between leaves)
  qd::DevBuf<int> ts_meta;       // interchanges of a leaf
  qd::DevBuf<float> ts_lu;       // LU of a leaf's winners block (32 x 32)
  qd::DevBuf<float> ts_top;      // a leaf's original top rows (32 x 32)
  qd::DevBuf<unsigned char> trs_tmp;  // intermediate vector of the block triangular solves
  qd::DevBuf<int> lb_flags;      // error word, then per-block and per-part flags of the look-back solves
  qd::DevBuf<unsigned char> lb_part;  // look-back solves: partial sums of the parts of each block row
  qd::DevBuf<unsigned long long> lb_pub;  // FP32 look-back solves: (epoch << 32) | bits of each x value
  int lb_epoch = 0;              // value the flags take in the latest look-back solve
  qd::DevBuf<float> ts_linv;     // inverses of the leaves' unit lower blocks (32 x 32 each)
  qd::DevBuf<unsigned> growth;   // diagonal-pivot factorisations: largest |l| (float bits)
  bool diag_pivots = false;      // the held factors took their diagonal pivots
  qd::DevBuf<float> inv32;       // inverted diagonal blocks of L and U
  qd::DevBuf<double> inv64;
  double anorm = -1.0;           // ||A||_inf of the FP64 matrix
  qd::DevBuf<double> rowabs;     // row sums of |A| per 512-column slice, from A's conversion into W32
  int rowabs_nsl = 0;            // slices in rowabs for the held factors (0: none)
  ~qd_lu_state() {
    if (ev_cols) (void)hipEventDestroy(ev_cols);
    if (ev_panel) (void)hipEventDestroy(ev_panel);
    if (ev_top) (void)hipEventDestroy(ev_top);
    if (ev_u12) (void)hipEventDestroy(ev_u12);
    if (aux) (void)hipStreamDestroy(aux);
  }
};

namespace qd {

constexpr int kLeaf = 32;
constexpr int kColThreads = 256;
#ifndef QD_COLSTEP_ROWS
#define QD_COLSTEP_ROWS 64
#endif
constexpr int kColRowsPerBlock = QD_COLSTEP_ROWS;

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// layout conversion: col-major double A (lda) -> row-major T W (ldw)
template <typename T>
__global__ __launch_bounds__(256) void k_transpose_in(long long N, const double* __restrict__ A,
                                                      long long lda, T* __restrict__ W, long long ldw,
                                                      long long ct0 = 0) {
  __shared__ double tile[32][33];
  const long long c0 = ((long long)blockIdx.x + ct0) * 32, r0 = (long long)blockIdx.y * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    long long r = r0 + tx, c = c0 + k;
    tile[k][tx] = (r < N && c < N) ? A[c * lda + r] : 0.0;  // tile[col][row]
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    long long r = r0 + k, c = c0 + tx;
    if (r < N && c < N) W[r * ldw + c] = (T)tile[tx][k];
  }
}

// k_transpose_in over whole 512-column slices (kAbsTiles tiles of 32 columns
// per workgroup, from column tile ct0, a multiple of kAbsTiles), also writing
// each row's sum of |A| over the slice to rowabs[slice * N + row]: the
// refinement's ||A||_inf without a pass of its own over A.
constexpr int kAbsTiles = 16;
template <typename T>
__global__ __launch_bounds__(256) void k_transpose_abs(long long N, const double* __restrict__ A, long long lda,
                                                       T* __restrict__ W, long long ldw, long long ct0,
                                                       double* __restrict__ rowabs) {
  __shared__ double tile[32][33];
  __shared__ double red[8][33];
  const long long r0 = (long long)blockIdx.y * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 32 x 8
  const long long t0 = ct0 + (long long)blockIdx.x * kAbsTiles;
  double acc = 0.0;
  for (int t = 0; t < kAbsTiles; ++t) {
    const long long c0 = (t0 + t) * 32;
    if (c0 >= N) break;
    for (int k = ty; k < 32; k += 8) {
      const long long r = r0 + tx, c = c0 + k;
      const double v = (r < N && c < N) ? A[c * lda + r] : 0.0;
      tile[k][tx] = v;  // tile[col][row]
      acc += fabs(v);
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
      const long long r = r0 + k, c = c0 + tx;
      if (r < N && c < N) W[r * ldw + c] = (T)tile[tx][k];
    }
    __syncthreads();
  }
  red[ty][tx] = acc;
  __syncthreads();
  if (ty == 0 && r0 + tx < N) {
    double sum = 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j) sum += red[j][tx];
    rowabs[(t0 / kAbsTiles) * N + r0 + tx] = sum;
  }
}

// rowabs slices summed: out[r] = sum over slices of rowabs[slice * N + r].
__global__ void k_rowabs_sum(long long N, int nsl, const double* __restrict__ rowabs, double* __restrict__ out) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= N) return;
  double s = 0.0;
  for (int q = 0; q < nsl; ++q) s += rowabs[(long long)q * N + r];
  out[r] = s;
}

// row-major double W -> col-major double A (LAPACK output of dgesv_)
__global__ __launch_bounds__(256) void k_transpose_out(long long N, const double* __restrict__ W, long long ldw,
                                                       double* __restrict__ A, long long lda) {
  __shared__ double tile[32][33];
  const long long c0 = (long long)blockIdx.x * 32, r0 = (long long)blockIdx.y * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
  for (int k = ty; k < 32; k += 8) {
    long long r = r0 + k, c = c0 + tx;
    tile[k][tx] = (r < N && c < N) ? W[r * ldw + c] : 0.0;  // tile[row][col]
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    long long r = r0 + tx, c = c0 + k;
    if (r < N && c < N) A[c * lda + r] = tile[tx][k];
  }
}

// ---------------------------------------------------------------------------
// One column step of a leaf.  rows [c, N) of the grid:
//  - when prev >= 0: multipliers of column prev (= c-1) and its rank-1 update
//    on leaf columns [c, lend);
//  - when c < lend: pivot search in column c; the last block picks the pivot,
//    swaps rows c and p across panel columns [pstart, pend) and stores 1/pivot.
// 32 lanes cover one row's leaf segment (columns lstart + lane), so every row
// access is one contiguous 128/256-byte segment; 8 rows per pass.
template <typename T>
__global__ __launch_bounds__(kColThreads) void k_colstep(T* __restrict__ W, long long ld, long long N, long long c,
                                                         long long prev, long long lstart, long long lend,
                                                         long long pstart, long long pend, double* __restrict__ part,
                                                         int* __restrict__ pidx, int* __restrict__ state,
                                                         double* __restrict__ rpiv, int* __restrict__ ipiv) {
  __shared__ double sv[kColThreads];
  __shared__ int si[kColThreads];
  __shared__ int am_last;
  const int tid = threadIdx.x;
  const int t = tid & 31, slot = tid >> 5;
  const long long col = lstart + t;
  const bool incol = col < lend;
  double best = -1.0;
  int bidx = 0x7fffffff;
  const double rp = (prev >= 0) ? rpiv[0] : 0.0;
  const T pv = (prev >= 0 && incol && col > prev) ? W[prev * ld + col] : (T)0;
  const long long rb = c + (long long)blockIdx.x * kColRowsPerBlock;
  for (int q = 0; q < kColRowsPerBlock / 8; ++q) {
    const long long r = rb + q * 8 + slot;
    if (r >= N) break;
    T* row = W + r * ld;
    if (prev >= 0) {
      const T l = (T)((double)row[prev] * rp);
      if (incol && col >= prev) row[col] = (col == prev) ? l : row[col] - l * pv;
    }
    if (c < lend && col == c) {
      const double v = fabs((double)row[c]);
      if (v > best) {
        best = v;
        bidx = (int)r;
      }
    }
  }
  if (c >= lend) return;
  sv[tid] = best;
  si[tid] = bidx;
  __syncthreads();
  for (int s = kColThreads / 2; s > 0; s >>= 1) {
    if (tid < s) {
      const double ov = sv[tid + s];
      const int oi = si[tid + s];
      if (ov > sv[tid] || (ov == sv[tid] && oi < si[tid])) {
        sv[tid] = ov;
        si[tid] = oi;
      }
    }
    __syncthreads();
  }
  // Publish the block maximum, then count the block in.  One device-scope
  // fence per block (by the thread that publishes), after the barrier that
  // orders the block's row updates before it.
  __syncthreads();
  if (tid == 0) {
    part[blockIdx.x] = sv[0];
    pidx[blockIdx.x] = si[0];
#ifdef QD_COLSTEP_FENCE_ALL
  }
  __threadfence();
  __syncthreads();
  if (tid == 0) {
#else
    __threadfence();
#endif
    const int t = atomicAdd(&state[0], 1);
    am_last = (t == (int)gridDim.x - 1);
  }
  __syncthreads();
  if (!am_last) return;
  __threadfence();
  // last block: reduce the block maxima
  double b2 = -1.0;
  int i2 = 0x7fffffff;
  for (int k = tid; k < (int)gridDim.x; k += kColThreads) {
    const double ov = __hip_atomic_load(&part[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int oi = __hip_atomic_load(&pidx[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (ov > b2 || (ov == b2 && oi < i2)) {
      b2 = ov;
      i2 = oi;
    }
  }
  sv[tid] = b2;
  si[tid] = i2;
  __syncthreads();
  for (int s = kColThreads / 2; s > 0; s >>= 1) {
    if (tid < s) {
      const double ov = sv[tid + s];
      const int oi = si[tid + s];
      if (ov > sv[tid] || (ov == sv[tid] && oi < si[tid])) {
        sv[tid] = ov;
        si[tid] = oi;
      }
    }
    __syncthreads();
  }
  long long p = si[0];
  const double pmax = sv[0];
  if (pmax <= 0.0 || p >= N) p = c;  // zero column: no interchange (LAPACK keeps going)
  if (p != c) {
    T* a = W + c * ld;
    T* b = W + p * ld;
    for (long long j = pstart + tid; j < pend; j += kColThreads) {
      const T t = a[j];
      a[j] = b[j];
      b[j] = t;
    }
  }
  __syncthreads();
  if (tid == 0) {
    ipiv[c] = (int)p;
    const double piv = (double)W[c * ld + c];
    if (piv == 0.0) {
      rpiv[0] = 0.0;
      if (state[1] == 0) state[1] = (int)c + 1;
    } else {
      rpiv[0] = 1.0 / piv;
    }
    state[0] = 0;
  }
}

// Row interchanges ipiv[k0..k1) applied to columns [c0, c1) (sequentially per
// column, as LAPACK dlaswp).
template <typename T>
__global__ __launch_bounds__(256) void k_laswp(T* __restrict__ W, long long ld, long long c0, long long c1,
                                               int k0, int k1, const int* __restrict__ ipiv) {
  const long long c = c0 + (long long)blockIdx.x * 256 + threadIdx.x;
  if (c >= c1) return;
  for (int k = k0; k < k1; ++k) {
    const long long p = ipiv[k];
    if (p != k) {
      const T t = W[(long long)k * ld + c];
      W[(long long)k * ld + c] = W[p * ld + c];
      W[p * ld + c] = t;
    }
  }
}

// X = L^-1 B for a unit-lower L of <= 32 rows: rows [r0, r0+m) of W hold L in
// columns [r0, r0+m) and B in columns [c0, c1).
template <typename T>
__global__ __launch_bounds__(256) void k_trsm_small(T* __restrict__ W, long long ld, long long r0, int m,
                                                    long long c0, long long c1) {
  __shared__ T sL[32][33];
  {
    // the (at most 4) loads of this thread first, then the LDS stores
    T v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = threadIdx.x + 256 * q;
      v[q] = (T)0;
      if (i < m * m) v[q] = W[(r0 + i / m) * ld + r0 + i % m];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = threadIdx.x + 256 * q;
      if (i < m * m) sL[i / m][i % m] = v[q];
    }
  }
  __syncthreads();
  const long long c = c0 + (long long)blockIdx.x * 256 + threadIdx.x;
  if (c >= c1) return;
  T x[32];
#pragma unroll
  for (int i = 0; i < 32; ++i)
    if (i < m) x[i] = W[(r0 + i) * ld + c];
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    if (i < m) {
      T s = x[i];
#pragma unroll
      for (int k = 0; k < 32; ++k)
        if (k < i) s -= sL[i][k] * x[k];
      x[i] = s;
    }
  }
#pragma unroll
  for (int i = 0; i < 32; ++i)
    if (i < m) W[(r0 + i) * ld + c] = x[i];
}

// X = L^-1 B in one launch for the FP32 factorisation: B = W[r0:r0+m, c0:c1),
// L unit lower at W[r0:r0+m, r0:r0+m] (m a multiple of 32) whose 32 x 32
// diagonal blocks are the tournament leaves, with their inverses in linv
// (block at rows r0+32b: linv + (r0/32 + b) * 1024, row-major).  One workgroup
// per strip of kSW columns (8 or 32), held in LDS (m x kSW floats).  Block by
// block, the strip's 32 rows are solved with the block's inverse plus one
// correction (x += inv (b - L x), which brings the explicit inverse to the
// accuracy of substitution), and the rows below subtract L[rows, block] times
// them.  Those rows go in chunks of kStripChunk whose L segments are staged
// through LDS by coalesced loads, the next chunk's loads in flight while one
// is used.  All products are MFMA f32 16x16x4 (a 16-column tile; the narrow
// strips use half of it).
constexpr int kStripChunk = 128;

template <int kSW>
__global__ __launch_bounds__(256) void k_trsm_strip(float* __restrict__ W, long long ld, long long r0, int m,
                                                    long long c0, long long c1, const float* __restrict__ linv) {
  static_assert(kSW == 8 || kSW == 32, "strip widths");
  extern __shared__ float xs[];  // m x kSW
  // rows of 32 floats whose 16-byte pieces are XOR-swizzled by row (piece p
  // of row r at p ^ (r & 7)): an MFMA operand read takes two pieces of 16
  // rows at once, and the swizzle spreads those over the banks
  __shared__ __attribute__((aligned(16))) float sinv[32][32], sl[32][32];
  __shared__ float sb[32][kSW];
  constexpr int kSlots = 256 / kSW;
  constexpr int NT = (kSW + 15) / 16;  // MFMA column tiles
  const int tid = threadIdx.x, c = tid % kSW, slot = tid / kSW;
  const int wave = tid >> 6, lane = tid & 63, lr = lane & 15, lg = lane >> 4;
  const long long col = c0 + (long long)blockIdx.x * kSW + c;
  const bool cv = col < c1;
  {
    // the strip's rows: every load in flight before the LDS stores (m is at
    // most 1024 rows for the narrow strips, 256 for the wide ones)
    constexpr int kLd = (kSW == 8 ? 1024 : 256) / kSlots;
    float v[kLd];
#pragma unroll
    for (int q = 0; q < kLd; ++q) {
      const int r = slot + q * kSlots;
      v[q] = (cv && r < m) ? W[(r0 + r) * ld + col] : 0.0f;
    }
#pragma unroll
    for (int q = 0; q < kLd; ++q) {
      const int r = slot + q * kSlots;
      if (r < m) xs[r * kSW + c] = v[q];
    }
  }
  const float* inv0 = linv + (r0 / 32) * 1024;
  auto swz = [](int r) { return r & 7; };
  // a chunk's L rows as MFMA A operands, straight from memory (L is shared by
  // every workgroup and stays in L2): wave w's rows first + 32 w + 16 i + lr,
  // lane group g's k = 8 g .. 8 g + 7 of the diagonal block's columns
  auto load_a = [&](long long d0, int first, float (&av)[2][8]) {
#pragma unroll
    for (int i2 = 0; i2 < 2; ++i2) {
      const int row = min(first + 32 * wave + 16 * i2 + lr, m - 1);
      const float* src = W + (r0 + row) * ld + d0 + 8 * lg;
      const float4 p0 = *reinterpret_cast<const float4*>(src);
      const float4 p1 = *reinterpret_cast<const float4*>(src + 4);
      av[i2][0] = p0.x;
      av[i2][1] = p0.y;
      av[i2][2] = p0.z;
      av[i2][3] = p0.w;
      av[i2][4] = p1.x;
      av[i2][5] = p1.y;
      av[i2][6] = p1.z;
      av[i2][7] = p1.w;
    }
  };
  // the k's of an MFMA step are permuted (lane group g supplies k = 8 g + t
  // at step t), so a lane's A values for the 32 k's are pieces 2g and 2g+1
  // of its row
  auto a_row = [&](const float* rowp, int r, float (&av)[8]) {
    const float4 a0 = *reinterpret_cast<const float4*>(rowp + 4 * ((2 * lg) ^ swz(r)));
    const float4 a1 = *reinterpret_cast<const float4*>(rowp + 4 * ((2 * lg + 1) ^ swz(r)));
    av[0] = a0.x;
    av[1] = a0.y;
    av[2] = a0.z;
    av[3] = a0.w;
    av[4] = a1.x;
    av[5] = a1.y;
    av[6] = a1.z;
    av[7] = a1.w;
  };
  // a diagonal block's inverse and L, four elements of each per thread
  float pinv[4], pl[4];
  auto load_diag = [&](int i) {
    const long long d0 = r0 + 32 * i;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u, rr = e >> 5, kk = e & 31;
      pinv[u] = inv0[(long long)i * 1024 + e];
      pl[u] = (kk < rr) ? W[(d0 + rr) * ld + d0 + kk] : (kk == rr ? 1.0f : 0.0f);
    }
  };
  // diagonal-block products: wave w forms the output tile at rows 16 (w % 2),
  // columns 16 (w / 2) (the narrow strips have one column tile, so waves 2
  // and 3 idle)
  const int ti = wave & 1, tj = wave >> 1;
  const bool dwave = tj < NT;
  const int arow = 16 * ti + lr, bcol = 16 * tj + lr;
  const bool bvalid = bcol < kSW;
  auto mfma_blk = [&](const float (*mat)[32], const float* bsrc) {
    float av[8];
    a_row(mat[arow], arow, av);
    f4 acc = (f4){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 8; ++t)
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[t], bvalid ? bsrc[(8 * lg + t) * kSW + bcol] : 0.0f, acc, 0, 0,
                                                 0);
    return acc;
  };
  load_diag(0);
  for (int i = 0; i < m / 32; ++i) {
    const long long d0 = r0 + 32 * i;  // the diagonal block's first row and column
    const int below = 32 * (i + 1), nch = (m - below + kStripChunk - 1) / kStripChunk;
    float an[2][8];
    if (nch > 0) load_a(d0, below, an);  // in flight during the block solve
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u, rr = e >> 5, kk = e & 31;
      const int ks = (((kk >> 2) ^ swz(rr)) << 2) | (kk & 3);
      sinv[rr][ks] = pinv[u];
      sl[rr][ks] = pl[u];
    }
    if (i + 1 < m / 32) load_diag(i + 1);  // in flight until the next block
    __syncthreads();
    // x = inv b, then x += inv (b - L x): three products on the 32 x kSW
    // block.  The triangular factors hold zeros above the diagonal, so the
    // products are dense.
    float* xb = xs + (32 * i) * kSW;  // the block's rows of the strip
    f4 xv = (f4){0.0f, 0.0f, 0.0f, 0.0f};
    float bown[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (dwave) {
      if (bvalid)
#pragma unroll
        for (int q = 0; q < 4; ++q) bown[q] = xb[(16 * ti + 4 * lg + q) * kSW + bcol];
      xv = mfma_blk(sinv, xb);
      if (bvalid)
#pragma unroll
        for (int q = 0; q < 4; ++q) sb[16 * ti + 4 * lg + q][bcol] = xv[q];
    }
    __syncthreads();
    if (dwave) {
      const f4 lx = mfma_blk(sl, &sb[0][0]);
      if (bvalid)
#pragma unroll
        for (int q = 0; q < 4; ++q) xb[(16 * ti + 4 * lg + q) * kSW + bcol] = bown[q] - lx[q];  // residual
    }
    __syncthreads();
    f4 dx = (f4){0.0f, 0.0f, 0.0f, 0.0f};
    if (dwave) dx = mfma_blk(sinv, xb);
    __syncthreads();
    if (dwave && bvalid)
#pragma unroll
      for (int q = 0; q < 4; ++q) xb[(16 * ti + 4 * lg + q) * kSW + bcol] = xv[q] + dx[q];
    __syncthreads();
    // the rows below, chunk by chunk: X_r -= L[r, block i] X_i.  Wave w
    // updates chunk rows [32 w, 32 w + 32) x the strip's columns, which no
    // other wave touches, so the chunks need no barrier; its X_i values are
    // loaded once per block and the next chunk's L rows while one is used.
    float xbf[NT][8];
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int t = 0; t < 8; ++t)
        xbf[j][t] = (16 * j + lr < kSW) ? xb[(8 * lg + t) * kSW + 16 * j + lr] : 0.0f;
    for (int ch = 0; ch < nch; ++ch) {
      const int first = below + ch * kStripChunk;
      float av[2][8];
#pragma unroll
      for (int i2 = 0; i2 < 2; ++i2)
#pragma unroll
        for (int t = 0; t < 8; ++t) av[i2][t] = an[i2][t];
      if (ch + 1 < nch) load_a(d0, first + kStripChunk, an);
      f4 acc[2][NT];
#pragma unroll
      for (int i2 = 0; i2 < 2; ++i2)
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          acc[i2][j] = (f4){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
          for (int t = 0; t < 8; ++t)
            acc[i2][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i2][t], xbf[j][t], acc[i2][j], 0, 0, 0);
        }
#pragma unroll
      for (int i2 = 0; i2 < 2; ++i2)
#pragma unroll
        for (int j = 0; j < NT; ++j)
          if (16 * j + lr < kSW)
#pragma unroll
            for (int q4 = 0; q4 < 4; ++q4) {
              const int r = first + 32 * wave + 16 * i2 + 4 * lg + q4;
              if (r < m) xs[r * kSW + 16 * j + lr] -= acc[i2][j][q4];
            }
    }
  }
  if (cv)
    for (int r = slot; r < m; r += kSlots) W[(r0 + r) * ld + col] = xs[r * kSW + c];
}

// ---------------------------------------------------------------------------
// C[M x N] -= A[M x K] * B[K x N], row-major operands inside W.
// 128x128 tile per 256-thread workgroup, each wave a 64x64 block of 4x4 MFMA
// 16x16 tiles; K staged through LDS (A transposed so fragment reads are
// contiguous).
constexpr int GBM = 128, GBN = 128;

template <typename T>
struct MfmaTraits;
template <>
struct MfmaTraits<double> {
  typedef d4 acc_t;
  static constexpr int BK = 16;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // output row of register r in lane group lk (rows interleave by 4)
  static __device__ __forceinline__ int out_row(int lk, int r) { return lk + 4 * r; }
};
template <>
struct MfmaTraits<float> {
  typedef f4 acc_t;
  static constexpr int BK = 32;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int out_row(int lk, int r) { return 4 * lk + r; }
};

template <typename T>
__global__ __launch_bounds__(256) void k_gemm_sub(long long M, long long N, int K, const T* __restrict__ A,
                                                  long long lda, const T* __restrict__ B, long long ldb,
                                                  T* __restrict__ C, long long ldc) {
  typedef MfmaTraits<T> MT;
  constexpr int BK = MT::BK;
  constexpr int PAD = 4;
  __shared__ T sA[BK][GBM + PAD];
  __shared__ T sB[BK][GBN + PAD];
  const long long m0 = (long long)blockIdx.y * GBM, n0 = (long long)blockIdx.x * GBN;
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int wm = (wave / 2) * 64, wn = (wave % 2) * 64;
  const int lr = lane % 16, lk = lane / 16;
  typename MT::acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (typename MT::acc_t){0, 0, 0, 0};
  // Interior tiles load without per-element tests, so all loads of a K step
  // are in flight together; edge tiles take the guarded loads.
  const bool full_mn = (m0 + GBM <= M) && (n0 + GBN <= N);
  for (int k0 = 0; k0 < K; k0 += BK) {
    if (full_mn && k0 + BK <= K) {
      T va[GBM * BK / 256], vb[GBN * BK / 256];
#pragma unroll
      for (int i = 0; i < GBM * BK / 256; ++i) {
        const int idx = tid + 256 * i;
        va[i] = A[(m0 + idx / BK) * lda + k0 + idx % BK];
      }
#pragma unroll
      for (int i = 0; i < GBN * BK / 256; ++i) {
        const int idx = tid + 256 * i;
        vb[i] = B[(long long)(k0 + idx / GBN) * ldb + n0 + idx % GBN];
      }
#pragma unroll
      for (int i = 0; i < GBM * BK / 256; ++i) {
        const int idx = tid + 256 * i;
        sA[idx % BK][idx / BK] = va[i];
      }
#pragma unroll
      for (int i = 0; i < GBN * BK / 256; ++i) {
        const int idx = tid + 256 * i;
        sB[idx / GBN][idx % GBN] = vb[i];
      }
    } else {
      // edge tile: every load predicated, all issued before the LDS stores
      T va[GBM * BK / 256], vb[GBN * BK / 256];
#pragma unroll
      for (int i = 0; i < GBM * BK / 256; ++i) {
        const int idx = tid + 256 * i;
        const long long gr = m0 + idx / BK;
        const int gc = k0 + idx % BK;
        va[i] = (T)0;
        if (gr < M && gc < K) va[i] = A[gr * lda + gc];
      }
#pragma unroll
      for (int i = 0; i < GBN * BK / 256; ++i) {
        const int idx = tid + 256 * i;
        const int gr = k0 + idx / GBN;
        const long long gc = n0 + idx % GBN;
        vb[i] = (T)0;
        if (gr < K && gc < N) vb[i] = B[(long long)gr * ldb + gc];
      }
#pragma unroll
      for (int i = 0; i < GBM * BK / 256; ++i) {
        const int idx = tid + 256 * i;
        sA[idx % BK][idx / BK] = va[i];
      }
#pragma unroll
      for (int i = 0; i < GBN * BK / 256; ++i) {
        const int idx = tid + 256 * i;
        sB[idx / GBN][idx % GBN] = vb[i];
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = sA[kk + lk][wm + 16 * i + lr];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = sB[kk + lk][wn + 16 * j + lr];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = MT::mma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
  if (full_mn) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long long gc = n0 + wn + 16 * j + lr;
      T cv[4][4];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) cv[i][r] = C[(m0 + wm + 16 * i + MT::out_row(lk, r)) * ldc + gc];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) C[(m0 + wm + 16 * i + MT::out_row(lk, r)) * ldc + gc] = cv[i][r] - acc[i][j][r];
    }
    return;
  }
  // edge tile: predicated loads of the column group first, then the stores
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const long long gc = n0 + wn + 16 * j + lr;
    T cv[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long long gr = m0 + wm + 16 * i + MT::out_row(lk, r);
        cv[i][r] = (T)0;
        if (gr < M && gc < N) cv[i][r] = C[gr * ldc + gc];
      }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long long gr = m0 + wm + 16 * i + MT::out_row(lk, r);
        if (gr < M && gc < N) C[gr * ldc + gc] = cv[i][r] - acc[i][j][r];
      }
  }
}

// FP32 C[M x N] -= A[M x K] B[K x N] for the narrow updates inside a panel
// (N <= NW, NW = 32 or 64), MFMA f32 16x16x4.  128 rows per workgroup, 32 per
// wave.  The sum over k may take the k's in any order as long as A and B
// agree, so in each 32-wide k slice lane group g of an MFMA supplies k = 8 g +
// step: a lane's A values for the slice are then 8 contiguous floats of its
// row, read straight from memory as two float4 (the four lane groups cover the
// row's 128 bytes).  The B slice (32 x NW) is staged in LDS.  The result goes
// through LDS so that C is read and written as float4 along its rows.
template <int NW>
__global__ __launch_bounds__(256) void k_gemm_narrow(long long M, int N, int K, const float* __restrict__ A,
                                                     long long lda, const float* __restrict__ B, long long ldb,
                                                     float* __restrict__ C, long long ldc) {
  constexpr int NT = NW / 16;
  __shared__ float sB[32][NW + 1];
  __shared__ __attribute__((aligned(16))) float sC[4][32][NW + 4];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lr = lane & 15, lg = lane >> 4;
  const long long m0 = (long long)blockIdx.x * 128 + wave * 32;
  const bool avec = ((reinterpret_cast<uintptr_t>(A) & 15) == 0) && (lda % 4 == 0);
  f4 acc[2][NT];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = (f4){0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < K; k0 += 32) {
    float a[2][8];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const long long row = m0 + 16 * i + lr;
      const int kb = k0 + 8 * lg;
      if (row < M && avec && kb + 8 <= K) {
        const float4 p0 = *reinterpret_cast<const float4*>(A + row * lda + kb);
        const float4 p1 = *reinterpret_cast<const float4*>(A + row * lda + kb + 4);
        a[i][0] = p0.x;
        a[i][1] = p0.y;
        a[i][2] = p0.z;
        a[i][3] = p0.w;
        a[i][4] = p1.x;
        a[i][5] = p1.y;
        a[i][6] = p1.z;
        a[i][7] = p1.w;
      } else {
#pragma unroll
        for (int t = 0; t < 8; ++t) a[i][t] = (row < M && kb + t < K) ? A[row * lda + kb + t] : 0.0f;
      }
    }
    // this slice of B: every load in flight before the LDS stores
    float bv[32 * NW / 256];
#pragma unroll
    for (int q = 0; q < 32 * NW / 256; ++q) {
      const int e = tid + 256 * q, kr = e / NW, c = e % NW;
      bv[q] = (k0 + kr < K && c < N) ? B[(long long)(k0 + kr) * ldb + c] : 0.0f;
    }
    if (k0 > 0) __syncthreads();  // the previous slice of B is consumed
#pragma unroll
    for (int q = 0; q < 32 * NW / 256; ++q) {
      const int e = tid + 256 * q;
      sB[e / NW][e % NW] = bv[q];
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 8; ++t) {
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const float b = sB[8 * lg + t][16 * j + lr];
#pragma unroll
        for (int i = 0; i < 2; ++i) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][t], b, acc[i][j], 0, 0, 0);
      }
    }
  }
  // the wave's 32 x NW result into LDS (row 16 i + 4 lg + r, column 16 j + lr)
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) sC[wave][16 * i + 4 * lg + r][16 * j + lr] = acc[i][j][r];
  __syncthreads();
  const bool cvec = ((reinterpret_cast<uintptr_t>(C) & 15) == 0) && (ldc % 4 == 0);
  constexpr int L4 = NW / 4;  // float4 pieces per row
  for (int e = lane; e < 32 * L4; e += 64) {
    const int rr = e / L4, c4 = (e % L4) * 4;
    const long long row = m0 + rr;
    if (row >= M) continue;
    const float4 d = *reinterpret_cast<const float4*>(&sC[wave][rr][c4]);
    float* cp = C + row * ldc + c4;
    if (cvec && c4 + 4 <= N) {
      float4 cv = *reinterpret_cast<const float4*>(cp);
      cv.x -= d.x;
      cv.y -= d.y;
      cv.z -= d.z;
      cv.w -= d.w;
      *reinterpret_cast<float4*>(cp) = cv;
    } else {
      if (c4 < N) cp[0] -= d.x;
      if (c4 + 1 < N) cp[1] -= d.y;
      if (c4 + 2 < N) cp[2] -= d.z;
      if (c4 + 3 < N) cp[3] -= d.w;
    }
  }
}

// FP64 C[M x N] -= A[M x K] B[K x N] (row-major, inside W) with MFMA
// f64 16x16x4.  128 x 128 tile per 256-thread workgroup (4 waves of 64 x 64,
// 4 x 4 MFMA tiles), K steps of 16 through double-buffered LDS with register
// prefetch, two workgroups per CU, XCD-banded tile order.
constexpr int D_BM = 128, D_BN = 128, D_BK = 16;
template <int kRole>
__global__ __launch_bounds__(256, 2) void k_gemm64(long long M, long long N, int K, const double* __restrict__ A,
                                                   long long lda, const double* __restrict__ B, long long ldb,
                                                   double* __restrict__ C, long long ldc) {
  __shared__ double sA[2][D_BM][D_BK + 1];  // [m][k]
  __shared__ double sB[2][D_BK][D_BN + 2];  // [k][n]
  const unsigned total = gridDim.x * gridDim.y;
  const unsigned bid = blockIdx.y * gridDim.x + blockIdx.x;
  const unsigned per_xcd = total / 8;
  const unsigned vid = (bid < per_xcd * 8) ? (bid % 8) * per_xcd + bid / 8 : bid;
  const long long m0 = (long long)(vid / gridDim.x) * D_BM, n0 = (long long)(vid % gridDim.x) * D_BN;
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int wm = (wave / 2) * 64, wn = (wave % 2) * 64;
  const int lr = lane % 16, lk = lane / 16;
  d4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (d4){0, 0, 0, 0};
  // A tile: 128 rows x 16: thread -> row tid/2, 8 consecutive k at (tid%2)*8
  // B tile: 16 rows x 128: thread -> row tid/16, 8 consecutive n at (tid%16)*8
  const int ar = tid / 2, ak = (tid % 2) * 8;
  const int br = tid / 16, bn = (tid % 16) * 8;
  double ra[8], rb[8];
  auto gload = [&](int k0) {
    const long long gr = m0 + ar;
#pragma unroll
    for (int q = 0; q < 8; ++q) ra[q] = (gr < M && k0 + ak + q < K) ? A[gr * lda + k0 + ak + q] : 0.0;
    const int gk = k0 + br;
#pragma unroll
    for (int q = 0; q < 8; ++q) rb[q] = (gk < K && n0 + bn + q < N) ? B[(long long)gk * ldb + n0 + bn + q] : 0.0;
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 8; ++q) sA[buf][ar][ak + q] = ra[q];
#pragma unroll
    for (int q = 0; q < 8; ++q) sB[buf][br][bn + q] = rb[q];
  };
  gload(0);
  lstore(0);
  __syncthreads();
  const int nk = (K + D_BK - 1) / D_BK;
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) gload((kt + 1) * D_BK);
#pragma unroll
    for (int kk = 0; kk < D_BK; kk += 4) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = sA[cur][wm + 16 * i + lr][kk + lk];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = sB[cur][kk + lk][wn + 16 * j + lr];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (kt + 1 < nk) lstore(cur ^ 1);
    __syncthreads();
  }
  // f64 16x16x4 output rows interleave by 4.  Interior tiles update C with
  // no per-element test, so each column group's 16 loads are issued together
  // before one wait (a guarded element would serialise load, wait, store).
  if (m0 + D_BM <= M && n0 + D_BN <= N) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long long gc = n0 + wn + 16 * j + lr;
      double cv[4][4];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) cv[i][r] = C[(m0 + wm + 16 * i + lk + 4 * r) * ldc + gc];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) C[(m0 + wm + 16 * i + lk + 4 * r) * ldc + gc] = cv[i][r] - acc[i][j][r];
    }
    return;
  }
  // edge tile: predicated loads of the column group first, then the stores
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const long long gc = n0 + wn + 16 * j + lr;
    double cv[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long long gr = m0 + wm + 16 * i + lk + 4 * r;
        cv[i][r] = 0.0;
        if (gr < M && gc < N) cv[i][r] = C[gr * ldc + gc];
      }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long long gr = m0 + wm + 16 * i + lk + 4 * r;
        if (gr < M && gc < N) C[gr * ldc + gc] = cv[i][r] - acc[i][j][r];
      }
  }
}

// FP16 trailing update.  Operands are converted once per update: A16 = L21
// rounded to fp16 (|l| <= 1 under partial pivoting, no scale needed), B16 =
// U12 transposed to (N x K) and divided by the per-64-column maxima colscale.
// Both are K-contiguous, K padded to a multiple of 32 with zeros.
//
// A16: one thread per 8 halves; rows read as two float4 when vec (A and lda
// 16-byte aligned, K a multiple of 8).
__global__ __launch_bounds__(256) void k_cvt_a16(long long M, int K, int Kp, const float* __restrict__ A, long long lda,
                                                 _Float16* __restrict__ A16, int vec) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const int per = Kp / 8;
  if (idx >= M * per) return;
  const long long r = idx / per;
  const int k0 = (int)(idx % per) * 8;
  h8 v;
  if (vec && k0 + 8 <= K) {
    const float4 p0 = *reinterpret_cast<const float4*>(A + r * lda + k0);
    const float4 p1 = *reinterpret_cast<const float4*>(A + r * lda + k0 + 4);
    v[0] = (_Float16)p0.x;
    v[1] = (_Float16)p0.y;
    v[2] = (_Float16)p0.z;
    v[3] = (_Float16)p0.w;
    v[4] = (_Float16)p1.x;
    v[5] = (_Float16)p1.y;
    v[6] = (_Float16)p1.z;
    v[7] = (_Float16)p1.w;
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = (_Float16)((k0 + q < K) ? A[r * lda + k0 + q] : 0.0f);
  }
  *(h8*)(A16 + r * Kp + k0) = v;
}

// B16 and colscale in one pass per 64-column block of B (K x N, row-major):
// the block's max |b| (1 where it is 0) goes to colscale, then the block is
// transposed through LDS in 32-row slices and written as B16[n][k] = b / scale
// (rows k >= K zero).  The second read of the block mostly hits L2.  vec: B
// and ldb 16-byte aligned, so interior rows are read as float4.
__global__ __launch_bounds__(256) void k_scale_cvt_b16t(int K, long long N, int Kp, const float* __restrict__ B,
                                                        long long ldb, float* __restrict__ colscale,
                                                        _Float16* __restrict__ B16, int vec) {
  __shared__ float red[4];
  __shared__ float tile[32][65];
  const int tid = threadIdx.x;
  const long long c0 = (long long)blockIdx.x * 64;
  const int cw = (int)min(64LL, N - c0);  // columns of this block
  const bool fullw = vec && cw == 64;
  const int cq = (tid % 16) * 4, rq = tid / 16;  // 16 rows per pass, 4 columns per thread
  float m = 0.0f;
  if (fullw) {
#pragma unroll 8
    for (int r = rq; r < K; r += 16) {
      const float4 v = *reinterpret_cast<const float4*>(B + (long long)r * ldb + c0 + cq);
      m = fmaxf(m, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
    }
  } else {
    for (int r = rq; r < K; r += 16)
#pragma unroll
      for (int t = 0; t < 4; ++t)
        if (cq + t < cw) m = fmaxf(m, fabsf(B[(long long)r * ldb + c0 + cq + t]));
  }
  // block maximum: DPP-free wave reduction through shuffles, then 4 waves
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if ((tid & 63) == 0) red[tid >> 6] = m;
  __syncthreads();
  const float mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  const float scale = mx > 0.0f ? mx : 1.0f;
  if (tid == 0) colscale[blockIdx.x] = scale;
  // transpose: slice rows [k0, k0+32) into tile[k][n], then 64 rows of B16.
  // The next slice's loads are issued before this slice's LDS round trip,
  // whose barriers wait on LDS only (s_barrier after lgkmcnt(0)), so they
  // stay in flight.
  const int wn = tid / 4, wk = (tid % 4) * 8;  // writer: column n, 8 k's
  float xa[2][4];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int k = k0 + rq + 16 * h;
      xa[h][0] = xa[h][1] = xa[h][2] = xa[h][3] = 0.0f;
      if (k < K) {
        if (fullw) {
          const float4 v = *reinterpret_cast<const float4*>(B + (long long)k * ldb + c0 + cq);
          xa[h][0] = v.x;
          xa[h][1] = v.y;
          xa[h][2] = v.z;
          xa[h][3] = v.w;
        } else {
          const float* src = B + (long long)k * ldb + c0 + cq;
#pragma unroll
          for (int t = 0; t < 4; ++t)
            if (cq + t < cw) xa[h][t] = src[t];
        }
      }
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < Kp; k0 += 32) {
    float xc[2][4];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int t = 0; t < 4; ++t) xc[h][t] = xa[h][t];
    if (k0 + 32 < Kp) fetch(k0 + 32);
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int t = 0; t < 4; ++t) tile[rq + 16 * h][cq + t] = xc[h][t] / scale;
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    if (wn < cw) {
      h8 o;
#pragma unroll
      for (int q = 0; q < 8; ++q) o[q] = (_Float16)tile[wk + q][wn];
      *(h8*)(B16 + (c0 + wn) * Kp + k0 + wk) = o;
    }
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
  }
}

// C[M x N] (fp32, row-major) -= colscale[n/64] * (A16 B16^T).  256 x 128 tile
// per 256-thread workgroup, 4 waves of 128 x 64 (8 x 4 MFMA 16x16x32 tiles),
// K steps of 32 through double-buffered LDS with register prefetch.
// Each workgroup loops over tiles.  The tiles are split into 8 contiguous
// bands, one per XCD (workgroups are dispatched to the XCDs round-robin), so
// an XCD's L2 serves the A/B rows its tiles share.  A grid smaller than the
// tile count leaves CU capacity to work on another stream.
#ifndef QD_G16_BM
#define QD_G16_BM 256
#endif
#ifndef QD_G16_OCC
#define QD_G16_OCC 2
#endif
// tile of H_BM x 128: four waves of (H_BM / 2) x 64, kG16Mi MFMA tiles of 16
// rows each per wave
constexpr int H_BM = QD_G16_BM, H_BN = 128, H_BK = 32;
constexpr int kG16Mi = H_BM / 32;
#ifndef QD_G16_GROUP
#define QD_G16_GROUP 4
#endif
constexpr unsigned kG16Group = QD_G16_GROUP;
#ifndef QD_G16_STAGES
#define QD_G16_STAGES 2
#endif
// Operand tiles go straight from global memory to LDS (global_load_lds_dwordx4:
// lane l of an instruction writes 16 bytes at the instruction's LDS address +
// 16 l), kG16Stages K steps in flight.  A stage holds A (256 rows) and B (128
// rows) as 64-byte rows of four 16-byte chunks; chunk c of row r sits in slot
// c ^ g16_swz(r), which keeps the 16-lane fragment reads of one MFMA operand
// on distinct banks.
constexpr int kG16Stages = QD_G16_STAGES;
constexpr int kG16StageBytes = (H_BM + H_BN) * H_BK * 2;
constexpr int kG16Loads = H_BM / 64 + 2;  // load instructions per stage and thread
static_assert(2 * kG16Loads < 16, "vmcnt immediates below use the low four bits");
struct G16Smem {
  unsigned char stage[kG16Stages][kG16StageBytes];
};
static_assert(4 * 32 * 68 * sizeof(float) <= sizeof(G16Smem), "epilogue staging must fit in the operand stages");
static_assert(H_BM == 128 || H_BM == 256, "tile height");

__device__ __forceinline__ int g16_swz(int row) { return ((row >> 1) ^ (row >> 3)) & 3; }

typedef __attribute__((address_space(3))) void* lds_ptr_t;

// Issues one stage's loads: A rows m0.., B rows n0.., K columns [k0, k0+32).
// Rows past M (N) read row M-1 (N-1): they feed only outputs that are not
// stored.  H_BM/64 + 2 instructions per thread, each covering 16 rows of A or B.
__device__ __forceinline__ void g16_issue(unsigned char* __restrict__ st, long long m0, long long n0, long long M,
                                          long long N, int Kp, int k0, const _Float16* __restrict__ A16,
                                          const _Float16* __restrict__ B16, int wave, int lane) {
  const int rl = lane >> 2, cs = lane & 3;  // row within the instruction, chunk slot
#pragma unroll
  for (int q = 0; q < H_BM / 64; ++q) {
    const int row = (q * 4 + wave) * 16 + rl;
    const long long g = min(m0 + row, M - 1);
    const int kc = cs ^ g16_swz(row);
    __builtin_amdgcn_global_load_lds(A16 + g * Kp + k0 + 8 * kc, (lds_ptr_t)(st + (q * 4 + wave) * 1024), 16, 0, 0);
  }
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int row = (q * 4 + wave) * 16 + rl;
    const long long g = min(n0 + row, N - 1);
    const int kc = cs ^ g16_swz(row);
    __builtin_amdgcn_global_load_lds(B16 + g * Kp + k0 + 8 * kc,
                                     (lds_ptr_t)(st + H_BM * 64 + (q * 4 + wave) * 1024), 16, 0, 0);
  }
}

// FP16 copy of an update's result (row-major, row pitch ld, rows counted from
// the update's first row) and the word that takes the largest |result|.
struct Out16 {
  _Float16* h;
  long long ld;
  unsigned* growth;
};
// Running maximum of |v| that keeps NaN (unsigned order of the magnitudes'
// bit patterns: NaN above +inf).
__device__ __forceinline__ float out16_absmax(float g, float v) {
  return __uint_as_float(max(__float_as_uint(g), __float_as_uint(fabsf(v))));
}
// The wave's largest magnitude folded into *growth (whole waves call it).
__device__ __forceinline__ void out16_growth(float g, unsigned* growth) {
  unsigned b = __float_as_uint(g);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) b = max(b, (unsigned)__shfl_xor((int)b, off));
  if ((threadIdx.x & 63) == 0) atomicMax(growth, b);
}
__device__ __forceinline__ void out16_put(const Out16& o, long long r, long long c, float v, float& g) {
  o.h[r * o.ld + c] = (_Float16)v;
  g = out16_absmax(g, v);
}
__device__ __forceinline__ void out16_put4(const Out16& o, long long r, long long c, float4 v, float& g) {
  typedef _Float16 h4 __attribute__((ext_vector_type(4)));
  *reinterpret_cast<h4*>(o.h + r * o.ld + c) = (h4){(_Float16)v.x, (_Float16)v.y, (_Float16)v.z, (_Float16)v.w};
  g = out16_absmax(out16_absmax(out16_absmax(out16_absmax(g, v.x), v.y), v.z), v.w);
}

template <bool kOut16>
__device__ __forceinline__ void gemm16_tile(long long m0, long long n0, long long M, long long N, int Kp, int Kt,
                                            const _Float16* __restrict__ A16, const _Float16* __restrict__ B16,
                                            const float* __restrict__ colscale, float* __restrict__ C, long long ldc,
                                            bool rows16, G16Smem& sm, const Out16& o16, float& gmax) {
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int wm = (wave / 2) * (H_BM / 2), wn = (wave % 2) * 64;
  const int lr = lane % 16, lk = lane / 16;
  f4 acc[kG16Mi][4];
#pragma unroll
  for (int i = 0; i < kG16Mi; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (f4){0, 0, 0, 0};
  const int nk = Kt / H_BK;  // K steps of this tile (Kp is the operands' row pitch)
  // fragment offsets within a stage (bytes): A row wm+16i+lr, B row wn+16j+lr, chunk lk
  int offa[kG16Mi], offb[4];
#pragma unroll
  for (int i = 0; i < kG16Mi; ++i) {
    const int row = wm + 16 * i + lr;
    offa[i] = row * 64 + 16 * (lk ^ g16_swz(row));
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int row = wn + 16 * j + lr;
    offb[j] = H_BM * 64 + row * 64 + 16 * (lk ^ g16_swz(row));
  }
  // every wave is done with the stages (previous tile, epilogue staging)
  __syncthreads();
#pragma unroll
  for (int s = 0; s < kG16Stages - 1; ++s)
    if (s < nk) g16_issue(sm.stage[s], m0, n0, M, N, Kp, s * H_BK, A16, B16, wave, lane);
  for (int kt = 0; kt < nk; ++kt) {
    // this step's stage has landed once at most the later stages' loads are
    // outstanding (kG16Loads per stage and thread)
    const int later = min(nk - 1 - kt, kG16Stages - 2);
    if (later >= 2) __builtin_amdgcn_s_waitcnt(0x0f70 | (2 * kG16Loads));
    else if (later == 1) __builtin_amdgcn_s_waitcnt(0x0f70 | kG16Loads);
    else __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0)
    // A plain barrier: __syncthreads() would also wait for every load in
    // flight (its release fence), which is the prefetch itself.  The previous
    // step's LDS reads have completed (the MFMAs consumed them), so the stage
    // the loads below overwrite is free.
    asm volatile("s_barrier" ::: "memory");
    if (kt + kG16Stages - 1 < nk)
      g16_issue(sm.stage[(kt + kG16Stages - 1) % kG16Stages], m0, n0, M, N, Kp, (kt + kG16Stages - 1) * H_BK, A16,
                B16, wave, lane);
    const unsigned char* st = sm.stage[kt % kG16Stages];
    h8 b[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = *reinterpret_cast<const h8*>(st + offb[j]);
#pragma unroll
    for (int i = 0; i < kG16Mi; ++i) {
      const h8 a = *reinterpret_cast<const h8*>(st + offa[i]);
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b[j], acc[i][j], 0, 0, 0);
    }
  }
  __syncthreads();
  // Interior tiles with 16-byte aligned C rows: each wave stages its
  // (H_BM/2) x 64 result through LDS (the operand buffers are free now) in
  // slices of 32 rows, then reads it back row-contiguously, so that C is read and
  // written in float4s covering 256 contiguous bytes per 16 lanes (the MFMA
  // layout alone gives 64-byte pieces).  The C accesses are non-temporal: C
  // is not touched again in the launch, and kept out of L2 it leaves room for
  // the operand tiles that neighbouring workgroups share.  The wave's 64
  // columns share one column scale.
  if (rows16 && m0 + H_BM <= M && n0 + H_BN <= N) {
    constexpr int kPitch = 68;  // floats per staged row
    float* stage = reinterpret_cast<float*>(&sm) + wave * (32 * kPitch);
    const float s = colscale[(n0 + wn) >> 6];
    // An opaque copy of ldc keeps the row-offset products inside the epilogue:
    // hoisted out of the tile loop they would stay live across the MFMA loop.
    long long ld = ldc;
    asm volatile("" : "+s"(ld));
    const int c4 = lane & 15;
    float* cbase = C + (m0 + wm + (lane >> 4)) * ld + n0 + wn + 4 * c4;
#pragma unroll
    for (int q = 0; q < kG16Mi / 2; ++q) {
#pragma unroll
      for (int ii = 0; ii < 2; ++ii)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) stage[(16 * ii + 4 * lk + r) * kPitch + 16 * j + lr] = acc[2 * q + ii][j][r];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        float4 cv[4], av[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int row = (lane >> 4) + 4 * (4 * h + t);
          {
            const f4 v = __builtin_nontemporal_load(reinterpret_cast<const f4*>(cbase + (32 * q + 4 * (4 * h + t)) * ld));
            cv[t] = make_float4(v[0], v[1], v[2], v[3]);
          }
          av[t] = *reinterpret_cast<const float4*>(stage + row * kPitch + 4 * c4);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          float4 o;
          o.x = cv[t].x - s * av[t].x;
          o.y = cv[t].y - s * av[t].y;
          o.z = cv[t].z - s * av[t].z;
          o.w = cv[t].w - s * av[t].w;
          __builtin_nontemporal_store((f4){o.x, o.y, o.z, o.w},
                                      reinterpret_cast<f4*>(cbase + (32 * q + 4 * (4 * h + t)) * ld));
          if constexpr (kOut16) {
            const long long orow = m0 + wm + (lane >> 4) + 32 * q + 4 * (4 * h + t);
            out16_put4(o16, orow, n0 + wn + 4 * c4, o, gmax);
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    return;
  }
  // Interior tiles otherwise: no per-element test, so each column group's 32
  // loads are issued together before one wait.
  if (m0 + H_BM <= M && n0 + H_BN <= N) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long long gc = n0 + wn + 16 * j + lr;
      const float s = colscale[gc >> 6];
      float cv[kG16Mi][4];
#pragma unroll
      for (int i = 0; i < kG16Mi; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) cv[i][r] = C[(m0 + wm + 16 * i + 4 * lk + r) * ldc + gc];
#pragma unroll
      for (int i = 0; i < kG16Mi; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float o = cv[i][r] - s * acc[i][j][r];
          C[(m0 + wm + 16 * i + 4 * lk + r) * ldc + gc] = o;
          if constexpr (kOut16) out16_put(o16, m0 + wm + 16 * i + 4 * lk + r, gc, o, gmax);
        }
    }
    return;
  }
  // edge tile: predicated loads of the column group first, then the stores
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const long long gc = n0 + wn + 16 * j + lr;
    const float s = (gc < N) ? colscale[gc >> 6] : 0.0f;
    float cv[kG16Mi][4];
#pragma unroll
    for (int i = 0; i < kG16Mi; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long long gr = m0 + wm + 16 * i + 4 * lk + r;
        cv[i][r] = 0.0f;
        if (gr < M && gc < N) cv[i][r] = C[gr * ldc + gc];
      }
#pragma unroll
    for (int i = 0; i < kG16Mi; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long long gr = m0 + wm + 16 * i + 4 * lk + r;
        if (gr < M && gc < N) {
          const float o = cv[i][r] - s * acc[i][j][r];
          C[gr * ldc + gc] = o;
          if constexpr (kOut16) out16_put(o16, gr, gc, o, gmax);
        }
      }
  }
}

// rows16: C and ldc keep every row start 16-byte aligned.  kRole 2 also
// writes the result as FP16 into o16 and its largest |c| into o16.growth.
// ktri: 0 = the whole K range; 1 = B16[n][k] is zero for k >= n (a tile
// runs K below its last column), 2 = A16[m][k] is zero for k >= m (K below
// its last row).
__device__ __forceinline__ int tile_kext(int Kp, int ktri, long long m0, long long n0, int bm, int bn) {
  if (ktri == 1) return (int)min((long long)Kp, n0 + bn);
  if (ktri == 2) return (int)min((long long)Kp, m0 + bm);
  return Kp;
}

template <int kRole>
__global__ __launch_bounds__(256, QD_G16_OCC) void k_gemm16(long long M, long long N, int Kp,
                                                            const _Float16* __restrict__ A16,
                                                const _Float16* __restrict__ B16, const float* __restrict__ colscale,
                                                float* __restrict__ C, long long ldc, int rows16, Out16 o16,
                                                int ktri) {
  __shared__ __attribute__((aligned(16))) G16Smem sm;
  float gmax = 0.0f;
  const unsigned tiles_n = (unsigned)((N + H_BN - 1) / H_BN);
  const unsigned tiles_m = (unsigned)((M + H_BM - 1) / H_BM);
  const unsigned total = tiles_m * tiles_n;
  const unsigned per_xcd = (total + 7) / 8;
  const unsigned xcd = blockIdx.x % 8, lw = blockIdx.x / 8, wx = (gridDim.x + 7 - xcd) / 8;
  for (unsigned i = lw; i < per_xcd; i += wx) {
    const unsigned vid = xcd * per_xcd + i;
    if (vid >= total) break;
    // grouped order: kG16Group row bands sweep the column tiles together, so
    // a B tile is fetched once for the group while the group's A rows stay
    // in L2 (row-band order re-streams all of B for every row band)
    const unsigned grp = vid / (kG16Group * tiles_n);
    const unsigned first_m = grp * kG16Group;
    const unsigned gsz = min(tiles_m - first_m, (unsigned)kG16Group);
    const unsigned in_grp = vid - grp * kG16Group * tiles_n;
    const unsigned tm = first_m + in_grp % gsz, tn = in_grp / gsz;
    const long long m0 = (long long)tm * H_BM, n0 = (long long)tn * H_BN;
    gemm16_tile<kRole == 2>(m0, n0, M, N, Kp, tile_kext(Kp, ktri, m0, n0, H_BM, H_BN), A16, B16, colscale, C, ldc,
                            rows16 != 0, sm, o16, gmax);
  }
  if constexpr (kRole == 2) out16_growth(gmax, o16.growth);
}

// ---------------------------------------------------------------------------
// Wide-tile FP16 update for the large launches: the same operation as
// k_gemm16 (C -= colscale[n/64] * (A16 B16^T), FP32 C row-major), with one
// 256 x 256 tile per workgroup of 4 waves, one wave per SIMD, each wave
// 128 x 128 as 4 x 4 blocks of MFMA f32 32x32x16 f16 (256 accumulators in
// AGPRs; 32x32x16 keeps one wave per SIMD at the full MFMA rate, 16x16x32
// does not).  K steps of 64 through two LDS stages of 64 KB, A and B as
// 128-byte rows, filled by global_load_lds: each instruction moves 8 whole
// rows (64-byte row pieces, as 32-wide stages give, move at half the rate
// into LDS).  The 16-byte chunks of a row are XOR-swizzled by row so that the
// fragment reads of 8 consecutive rows hit distinct banks.  Fragments are
// read one 32-wide half step ahead of their MFMAs.  The epilogue stages C
// through the free stages with global_load_lds as well (64-row quarters, two
// in flight, rows with their 16-byte units XOR-swizzled so that the two
// half-waves of a read hit different banks): the loads in flight hold no
// registers, which gives each CU ~2.4x the C bandwidth of register loads (a
// CU's epilogue is bound by its bytes in flight, not by HBM, until every CU
// is in it).  Kp must be a multiple of 64.
namespace w16 {
constexpr int BM = 256, BN = 256, BK = 64;
constexpr int kStage = (BM + BN) * BK * 2;  // 64 KB
constexpr int kLoads = (BM + BN) / 32;      // global_load_lds per thread per stage (8 rows each)
typedef float f16v __attribute__((ext_vector_type(16)));

struct Smem {
  unsigned char st[2][kStage];
};
static_assert(sizeof(Smem) == 2 * 65536, "the epilogue's two C quarters fill the stages");

// chunk c (0..7) of row r sits in slot c ^ swz(r): distinct for 8 consecutive
// rows and for each parity class of 16; period 32 rows, so the four 32-row
// blocks of a wave share their fragment offsets
__device__ __forceinline__ int swz(int row) { return (row & 7) ^ ((row >> 3) & 3); }

template <int N>
__device__ __forceinline__ void wait_vm() {
  static_assert(N >= 0 && N < 64, "vmcnt");
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}
__device__ __forceinline__ void wait_lgkm0() { __builtin_amdgcn_s_waitcnt(0xc07f); }

struct Frag {
  h8 a[2][4], b[2][4];  // [16-wide k substep of the half step][32-row block]
};

// fragments of half H (k 32H .. 32H+31) of a stage: substeps 2H, 2H+1
template <int H>
__device__ __forceinline__ void read_frags(const unsigned char* st, const int (&offa)[4], const int (&offb)[4],
                                           Frag& f) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    f.a[0][i] = *reinterpret_cast<const h8*>(st + offa[2 * H] + i * 32 * 128);
    f.a[1][i] = *reinterpret_cast<const h8*>(st + offa[2 * H + 1] + i * 32 * 128);
    f.b[0][i] = *reinterpret_cast<const h8*>(st + offb[2 * H] + i * 32 * 128);
    f.b[1][i] = *reinterpret_cast<const h8*>(st + offb[2 * H + 1] + i * 32 * 128);
  }
}

// Lane l of instruction q covers row (q*4 + wave)*8 + l/8 of A (and of B),
// 16-byte slot l%8; its offsets are 32-bit element offsets from A16 / B16
// (rows past M or N read row M-1 or N-1: they feed only outputs not stored).
struct Loader {
  const _Float16* A16;
  const _Float16* B16;
  unsigned oa[8], ob[8];
  __device__ __forceinline__ void issue(unsigned char* st, int kb, int wave) const {
    const _Float16* a = A16 + kb * BK;
    const _Float16* b = B16 + kb * BK;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      __builtin_amdgcn_global_load_lds(a + oa[q], (lds_ptr_t)(st + (q * 4 + wave) * 1024), 16, 0, 0);
#pragma unroll
    for (int q = 0; q < 8; ++q)
      __builtin_amdgcn_global_load_lds(b + ob[q], (lds_ptr_t)(st + ((q + 8) * 4 + wave) * 1024), 16, 0, 0);
  }
  // load Q (0..15) of issue() alone
  template <int Q>
  __device__ __forceinline__ void issue_one(unsigned char* st, int kb, int wave) const {
    if constexpr (Q < 8)
      __builtin_amdgcn_global_load_lds(A16 + kb * BK + oa[Q], (lds_ptr_t)(st + (Q * 4 + wave) * 1024), 16, 0, 0);
    else
      __builtin_amdgcn_global_load_lds(B16 + kb * BK + ob[Q - 8], (lds_ptr_t)(st + (Q * 4 + wave) * 1024), 16, 0, 0);
  }
};

template <int... Ts, class F>
__device__ __forceinline__ void seq_each(std::integer_sequence<int, Ts...>, F&& f) {
  (f(std::integral_constant<int, Ts>{}), ...);
}

// fragment U (0..15) of half H of a stage, in the order the MFMAs use them:
// substep U/8, A (U/4 even) or B, 32-row block U%4
template <int H, int U>
__device__ __forceinline__ void read_frag(const unsigned char* st, const int (&offa)[4], const int (&offb)[4],
                                          Frag& f) {
  constexpr int s = U / 8, i = U % 4;
  if constexpr ((U / 4) % 2 == 0) f.a[s][i] = *reinterpret_cast<const h8*>(st + offa[2 * H + s] + i * 32 * 128);
  else f.b[s][i] = *reinterpret_cast<const h8*>(st + offb[2 * H + s] + i * 32 * 128);
}

// Half step (stage k, half H) on the stage in slot SL.  Stage k's slot takes
// stage k+2 (kLoad 1: it exists, 2: only if it does) once every wave has read
// its second half, from the middle of H = 0 on; at H = 1 stage k+1 must have
// landed (its first half is read from now on).  The instructions go in a
// fixed order, one MFMA per group: the next half's 16 fragment reads beside
// the first 16 MFMAs (they have landed long before the next half needs them),
// and the 16 operand loads one per two MFMAs over the last 16 MFMAs of H = 0
// and the first 16 of H = 1.  Issued together after a barrier, the 64 loads
// of a CU's four waves queue at the address unit and hold each wave's
// in-order issue, MFMAs included (trailing-update replay 944 -> 1032 TFLOP/s
// with both changes).
template <int SL, int H, int kLoad>
__device__ __forceinline__ void hstep(int k, int nk, Smem& sm, const Loader& ld, int wave, const int (&offa)[4],
                                      const int (&offb)[4], Frag& cur, Frag& nxt, f16v (&acc)[4][4]) {
  const bool load = kLoad == 1 || k + 2 < nk;
  if constexpr (H == 1) {
    // stage k+1 has landed once at most the 8 loads of stage k+2 issued in
    // the first half are in flight
    if (kLoad == 1 || load) wait_vm<8>();
    else wait_vm<0>();
    asm volatile("s_barrier" ::: "memory");
  }
  // past the last stage the reads are stale and never used
  const unsigned char* rs = H == 1 ? sm.st[SL ^ 1] : sm.st[SL];
  auto one = [&](auto tc) {
    constexpr int t = decltype(tc)::value;
    constexpr int s = t / 16, i = (t / 4) % 4, j = t % 4;
    if constexpr (H == 0 && t == 17) {
      if (kLoad == 1 || load) {
        // every wave has read stage k (slot SL): it takes stage k+2
        wait_lgkm0();
        asm volatile("s_barrier" ::: "memory");
      }
    }
    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(cur.a[s][i], cur.b[s][j], acc[i][j], 0, 0, 0);
    if constexpr (t < 16) read_frag<H == 1 ? 0 : 1, t>(rs, offa, offb, nxt);
    if constexpr (H == 0 && t >= 17 && t % 2 == 1) {
      if (kLoad == 1 || load) ld.issue_one<(t - 17) / 2>(sm.st[SL], k + 2, wave);
    }
    if constexpr (H == 1 && t < 16 && t % 2 == 1) {
      if (kLoad == 1 || load) ld.issue_one<8 + t / 2>(sm.st[SL], k + 2, wave);
    }
    __builtin_amdgcn_sched_barrier(0);
  };
  seq_each(std::make_integer_sequence<int, 32>{}, one);
}

template <int kLoad, int... Us>
__device__ __forceinline__ void hsteps(std::integer_sequence<int, Us...>, int hs0, int nk, Smem& sm, const Loader& ld,
                                       int wave, const int (&offa)[4], const int (&offb)[4], Frag& f0, Frag& f1,
                                       f16v (&acc)[4][4]) {
  auto one = [&](auto uc) {
    constexpr int u = decltype(uc)::value;
    const int hs = hs0 + u;
    if (kLoad == 1 || hs < 2 * nk) {
      if constexpr (u % 2 == 0) hstep<(u / 2) % 2, u % 2, kLoad>(hs / 2, nk, sm, ld, wave, offa, offb, f0, f1, acc);
      else hstep<(u / 2) % 2, u % 2, kLoad>(hs / 2, nk, sm, ld, wave, offa, offb, f1, f0, acc);
    }
  };
  (one(std::integral_constant<int, Us>{}), ...);
}

template <bool kOut16>
__device__ __forceinline__ void tile(long long m0, long long n0, long long M, long long N, int Kp, int Kt,
                                     const _Float16* __restrict__ A16, const _Float16* __restrict__ B16,
                                     const float* __restrict__ colscale, float* __restrict__ C, long long ldc,
                                     bool rows16, Smem& sm, const Out16& o16, float& gmax) {
  const int tid = threadIdx.x, lane = tid % 64;
  const int wave = __builtin_amdgcn_readfirstlane(tid / 64);  // wave-uniform: LDS addresses in SGPRs
  const int wm = (wave / 2) * 128, wn = (wave % 2) * 128;
  Loader ld;
  ld.A16 = A16;
  ld.B16 = B16;
  {
    const int rl = lane >> 3, cs = lane & 7;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int row = (q * 4 + wave) * 8 + rl;
      const int kc = cs ^ swz(row);
      ld.oa[q] = (unsigned)(min(m0 + row, M - 1) * Kp + 8 * kc);
      ld.ob[q] = (unsigned)(min(n0 + row, N - 1) * Kp + 8 * kc);
    }
  }
  // MFMA 32x32x16 operands: lane l supplies row l%32 of a 32-row block, k
  // chunk 2s + l/32 of the stage's eight 8-wide chunks at substep s
  const int l32 = lane & 31, lh = lane >> 5;
  const int ra = wm + l32, rb = wn + l32;
  int offa[4], offb[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    offa[s] = ra * 128 + 16 * ((2 * s + lh) ^ swz(ra));
    offb[s] = BM * 128 + rb * 128 + 16 * ((2 * s + lh) ^ swz(rb));
  }
  f16v acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
  const int nk = Kt / BK;  // K steps of this tile (Kp is the operands' row pitch)
  __syncthreads();  // every wave is done with the previous tile's stages
  ld.issue(sm.st[0], 0, wave);
  if (nk > 1) {
    ld.issue(sm.st[1], 1, wave);
    wait_vm<kLoads>();
  } else {
    wait_vm<0>();
  }
  asm volatile("s_barrier" ::: "memory");
  Frag f0, f1;
  read_frags<0>(sm.st[0], offa, offb, f0);
  // stage k sits in slot k & 1; the hot loop's 4 half steps cover stages k,
  // k+1 (slots 0, 1) and load stages k+2, k+3
  int hs = 0;
  for (; hs / 2 + 3 < nk; hs += 4)
    hsteps<1>(std::make_integer_sequence<int, 4>{}, hs, nk, sm, ld, wave, offa, offb, f0, f1, acc);
  // the rest: at most 3 stages (hs is a multiple of 4: slot 0, fragments f0)
  hsteps<2>(std::make_integer_sequence<int, 6>{}, hs, nk, sm, ld, wave, offa, offb, f0, f1, acc);
  // Epilogue.  acc[i][j][r] is C row wm + 32 i + 8 (r/4) + 4 lh + r%4, column
  // wn + 32 j + l32 of the tile; a column block of 32 shares its scale.
  if (m0 + BM <= M && n0 + BN <= N && rows16) {
    unsigned char* lds = &sm.st[0][0];
    // quarter q: tile rows 32q .. 32q+31 of each wave row (wm = 0, 128), i.e.
    // acc block row q of every wave; quarter row lr = (wm/128)*32 + row in block
    auto issue_q = [&](int q) {
      unsigned char* region = lds + (q & 1) * 65536;
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const int lr = wave * 16 + t;
        const long long grow = m0 + (lr / 32) * 128 + 32 * q + (lr % 32);
        const int u = lane ^ (8 * ((lr >> 2) & 1));
        // cache policy 2: non-temporal (C is not read again in the launch)
        __builtin_amdgcn_global_load_lds(C + grow * ldc + n0 + 4 * u, (lds_ptr_t)(region + lr * 1024), 16, 0, 2);
      }
    };
    float sc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) sc[j] = colscale[(n0 + wn + 32 * j) >> 6];
    float* cbase = C + (m0 + wm + 4 * lh) * ldc + n0 + wn + l32;
    __syncthreads();  // every wave is done with the operand stages
    issue_q(0);
    issue_q(1);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      // Quarter q has landed.  At q = 0 the next quarter's 16 loads may be in
      // flight; later also this wave's stores of the previous quarter (up to
      // 80, in order, of which 63 can be waited for).
      if (q == 0) wait_vm<16>();
      else wait_vm<63>();
      asm volatile("s_barrier" ::: "memory");
      const unsigned char* region = lds + (q & 1) * 65536;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int lr = (wm / 128) * 32 + 8 * (r / 4) + 4 * lh + (r % 4);
        float* p = cbase + (32 * q + 8 * (r / 4) + (r % 4)) * ldc;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int col = wn + 32 * j + l32;
          const float cv = *reinterpret_cast<const float*>(region + lr * 1024 +
                                                          16 * ((col >> 2) ^ (8 * ((lr >> 2) & 1))) + 4 * (col & 3));
          const float o = cv - sc[j] * acc[q][j][r];
          __builtin_nontemporal_store(o, p + 32 * j);
          if constexpr (kOut16)
            out16_put(o16, m0 + wm + 4 * lh + 32 * q + 8 * (r / 4) + (r % 4), n0 + col, o, gmax);
        }
      }
      if (q + 2 < 4) {
        // every wave has read quarter q: its region takes quarter q + 2
        wait_lgkm0();
        asm volatile("s_barrier" ::: "memory");
        issue_q(q + 2);
      }
    }
    return;
  }
  // edge tiles (and C rows that are not 16-byte aligned): predicated
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float cv[4][16];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long long col = n0 + wn + 32 * j + l32;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const long long row = m0 + wm + 32 * i + 8 * (r / 4) + 4 * lh + (r % 4);
        cv[j][r] = (row < M && col < N) ? C[row * ldc + col] : 0.0f;
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long long col = n0 + wn + 32 * j + l32;
      const long long cb = n0 + wn + 32 * j;
      const float s = (cb < N) ? colscale[cb >> 6] : 0.0f;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const long long row = m0 + wm + 32 * i + 8 * (r / 4) + 4 * lh + (r % 4);
        if (row < M && col < N) {
          const float o = cv[j][r] - s * acc[i][j][r];
          C[row * ldc + col] = o;
          if constexpr (kOut16) out16_put(o16, row, col, o, gmax);
        }
      }
    }
  }
}
}  // namespace w16

// One 256 x 256 tile per workgroup; the tiles go to the XCDs in 8 contiguous
// bands and, within a band, 4 row bands sweep the column tiles together (as
// in k_gemm16).
template <int kRole>
__global__ __launch_bounds__(256, 1) void k_gemm16w(long long M, long long N, int Kp, const _Float16* __restrict__ A16,
                                                     const _Float16* __restrict__ B16,
                                                     const float* __restrict__ colscale, float* __restrict__ C,
                                                     long long ldc, int rows16, Out16 o16, int ktri) {
  __shared__ __attribute__((aligned(16))) w16::Smem sm;
  float gmax = 0.0f;
  const unsigned tiles_n = (unsigned)((N + w16::BN - 1) / w16::BN);
  const unsigned tiles_m = (unsigned)((M + w16::BM - 1) / w16::BM);
  const unsigned total = tiles_m * tiles_n;
  const unsigned per_xcd = (total + 7) / 8;
  const unsigned xcd = blockIdx.x % 8, lw = blockIdx.x / 8, wx = (gridDim.x + 7 - xcd) / 8;
  for (unsigned i = lw; i < per_xcd; i += wx) {
    const unsigned vid = xcd * per_xcd + i;
    if (vid >= total) break;
    const unsigned grp = vid / (kG16Group * tiles_n);
    const unsigned first_m = grp * kG16Group;
    const unsigned gsz = min(tiles_m - first_m, (unsigned)kG16Group);
    const unsigned in_grp = vid - grp * kG16Group * tiles_n;
    const unsigned tm = first_m + in_grp % gsz, tn = in_grp / gsz;
    const long long m0 = (long long)tm * w16::BM, n0 = (long long)tn * w16::BN;
    w16::tile<kRole == 2>(m0, n0, M, N, Kp, tile_kext(Kp, ktri, m0, n0, w16::BM, w16::BN), A16, B16, colscale, C,
                          ldc, rows16 != 0, sm, o16, gmax);
  }
  if constexpr (kRole == 2) out16_growth(gmax, o16.growth);
}

}  // namespace qd

#include "qd_tslu.h"

namespace qd {

// ---------------------------------------------------------------------------
// Diagonal-block panels.  With diagonal pivots a panel [j, j+jb) needs no
// search over the rows below it: only its jb x jb diagonal block is factored
// as a chain of small launches (D = L11 U11), and the multipliers below follow
// from one product, L21 = A21 U11^-1.  U11^-1 is formed explicitly by
// recursive doubling on 64-column blocks (the inverse of [U1 U12; 0 U2] is
// [Y1, -Y1 U12 Y2; 0, Y2]), in FP32, padded to npad = 64 * 2^k with the
// identity.  The product runs on the FP16 update kernels with both operands
// scaled to O(1) entries through the diagonal of U11: A16 = -A21 diag(1/u_kk)
// and B = diag(u_kk) U11^-1 (the inverse of U11 with its columns divided by
// their diagonal entries: unit upper triangular), so that C -= colscale *
// (A16 B16^T) on a cleared C leaves A21 U11^-1.

// npad of a diagonal block of jb columns.
static inline int binv_pad(int jb) {
  int p = 64;
  while (p < jb) p *= 2;
  return p;
}

// ---- the diagonal block in 64-column steps ----------------------------------
// The jb x jb diagonal block (jb a multiple of 64, origin (c0, c0) in W) in
// 64 x 64 tiles, one launch per step k (k_blk_step): the workgroup of
// trailing tile (i, j), i, j > k, forms L_ik = A_ik U_kk^-1 and U_kj = L_kk^-1
// A_kj itself, from the inverses of tile k's factors and from copies of tile
// column and row k that the previous step left in colbuf / rowbuf, then
// subtracts L_ik U_kj from its tile.  Workgroups of tile column k+1 also
// store L_ik into W and their result into the next step's colbuf, those of
// tile row k+1 U_kj and the next rowbuf; workgroup 0, tile (k+1, k+1),
// factors its result.  Step buffers alternate by the parity of k, so no
// workgroup reads what another writes in the same launch.  A launch is a few
// microseconds of work: the chain of jb / 64 launches replaces the recursive
// panel's four to five per 64 columns.

// Row R of the substitutions for column `lane` of a 64 x 64 triangular
// factor's inverse (sT rows: L below the diagonal, unit; U on and above it),
// unrolled through integer sequences so that y stays in registers.
template <int R>
__device__ __forceinline__ void blk_fwd_row(const float (*sT)[68], int lane, float (&y)[kDW]) {
  float acc = (R == lane) ? 1.0f : 0.0f;
#pragma unroll
  for (int m = 0; m < R; ++m) acc -= sT[R][m] * y[m];
  y[R] = acc;
}
template <int... Rs>
__device__ __forceinline__ void blk_fwd(std::integer_sequence<int, Rs...>, const float (*sT)[68], int lane,
                                        float (&y)[kDW]) {
  (blk_fwd_row<Rs>(sT, lane, y), ...);
}
template <int R>
__device__ __forceinline__ void blk_bwd_row(const float (*sT)[68], int lane, float (&y)[kDW]) {
  float acc = (R == lane) ? 1.0f : 0.0f;
#pragma unroll
  for (int m = R + 1; m < kDW; ++m) acc -= sT[R][m] * y[m];
  y[R] = acc / sT[R][R];
}
template <int... Rs>
__device__ __forceinline__ void blk_bwd(std::integer_sequence<int, Rs...>, const float (*sT)[68], int lane,
                                        float (&y)[kDW]) {
  (blk_bwd_row<kDW - 1 - Rs>(sT, lane, y), ...);
}

// Factor of the 64 x 64 tile held in sT (rows) at block tile t, by a
// 256-thread workgroup: wave 0 eliminates with register rows and readlane
// pivot-row broadcasts (the diagonal-pivot leaf's code), stores the LU rows
// into W, ipiv = identity, and folds its largest multiplier (a zero pivot
// counts as +inf) into *growth; then wave 1 forms L_tt^-1 (into inv, Z's
// diagonal block t, and its two 32 x 32 diagonal blocks into linv for the
// strip solves of U12 outside the block), wave 2 U_tt^-1 (into inv + 4096
// and Y's diagonal block t; Y and Z have row pitch npad), wave 3 dg[k] =
// u_kk, dg[npad + k] = -1 / u_kk.
__device__ __forceinline__ void blk_factor_tile(float (*sT)[68], float* __restrict__ W, long long ld, long long c0,
                                                int t, float* __restrict__ inv, float* __restrict__ Y,
                                                float* __restrict__ Z, int npad, float* __restrict__ dg,
                                                float* __restrict__ linv, int* __restrict__ ipiv,
                                                unsigned* __restrict__ growth) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long long g0 = c0 + 64LL * t;
  if (wave == 0) {
    float a[kDW];
#pragma unroll
    for (int j = 0; j < kDW; ++j) a[j] = sT[lane][j];
    float gmax = 0.0f;
    dw_all64(std::make_integer_sequence<int, kDW>{}, a, gmax);
    float* dst = W + (g0 + lane) * ld + g0;
#pragma unroll
    for (int j4 = 0; j4 < kDW / 4; ++j4)
      *reinterpret_cast<float4*>(dst + 4 * j4) = make_float4(a[4 * j4], a[4 * j4 + 1], a[4 * j4 + 2], a[4 * j4 + 3]);
#pragma unroll
    for (int j = 0; j < kDW; ++j) sT[lane][j] = a[j];
    ts_growth_max(gmax, growth);
    ipiv[g0 + lane] = (int)(g0 + lane);
  }
  __syncthreads();
  if (wave == 1) {
    // lane c: column c of L_tt^-1 by forward substitution (unit diagonal)
    float y[kDW];
    blk_fwd(std::make_integer_sequence<int, kDW>{}, sT, lane, y);
#pragma unroll
    for (int r = 0; r < kDW; ++r) inv[r * kDW + lane] = y[r];
    float* zo = Z + (64LL * t) * npad + 64LL * t + lane;
#pragma unroll
    for (int r = 0; r < kDW; ++r) zo[(long long)r * npad] = y[r];
    // the 32 x 32 diagonal blocks of L^-1 are the inverses of L's
    // (two guarded loops, so that y is indexed by constants only)
    const int cc = lane & 31;
    float* out = linv + (g0 / kTsW) * kTsW * kTsW;
    if (lane < 32) {
#pragma unroll
      for (int r = 0; r < 32; ++r) out[r * 32 + cc] = y[r];
    } else {
#pragma unroll
      for (int r = 0; r < 32; ++r) out[1024 + r * 32 + cc] = y[32 + r];
    }
  } else if (wave == 2) {
    // lane c: column c of U_tt^-1 by back substitution
    float y[kDW];
    blk_bwd(std::make_integer_sequence<int, kDW>{}, sT, lane, y);
#pragma unroll
    for (int r = 0; r < kDW; ++r) inv[kDW * kDW + r * kDW + lane] = y[r];
    float* out = Y + (64LL * t) * npad + 64LL * t + lane;
#pragma unroll
    for (int r = 0; r < kDW; ++r) out[(long long)r * npad] = y[r];
  } else if (wave == 3) {
    const float d = sT[lane][lane];
    dg[64 * t + lane] = d;
    dg[npad + 64 * t + lane] = -1.0f / d;
  }
}

// 64 x 64 tile (row pitch ld) into LDS rows, and transposed (sX[c][r]).
__device__ __forceinline__ void blk_load(const float* __restrict__ src, long long ld, float4 (&v)[4]) {
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int idx = threadIdx.x + 256 * p, r = idx / 16, c4 = (idx % 16) * 4;
    v[p] = *reinterpret_cast<const float4*>(src + r * ld + c4);
  }
}
__device__ __forceinline__ void blk_put(float (*sX)[68], const float4 (&v)[4]) {
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int idx = threadIdx.x + 256 * p, r = idx / 16, c4 = (idx % 16) * 4;
    *reinterpret_cast<float4*>(&sX[r][c4]) = v[p];
  }
}
__device__ __forceinline__ void blk_put_t(float (*sX)[68], const float4 (&v)[4]) {
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int idx = threadIdx.x + 256 * p, r = idx / 16, c4 = (idx % 16) * 4;
    sX[c4][r] = v[p].x;
    sX[c4 + 1][r] = v[p].y;
    sX[c4 + 2][r] = v[p].z;
    sX[c4 + 3][r] = v[p].w;
  }
}

// acc (this wave's 32 x 32 quarter at (wm, wn)) = sAt^T sB over k < 64, with
// sAt[k][m] and sB[k][n]; MFMA f32 16x16x4.
__device__ __forceinline__ void blk_mm(const float (*sAt)[68], const float (*sB)[68], int wm, int wn, int lr, int lk,
                                       f4 (&acc)[2][2]) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (f4){0, 0, 0, 0};
#pragma unroll
  for (int kk = 0; kk < kDW; kk += 4) {
    float a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i] = sAt[kk + lk][wm + 16 * i + lr];
#pragma unroll
    for (int j = 0; j < 2; ++j) b[j] = sB[kk + lk][wn + 16 * j + lr];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
  }
}

// acc += sAt^T sB (as blk_mm, without clearing acc).
__device__ __forceinline__ void blk_mm_acc(const float (*sAt)[68], const float (*sB)[68], int wm, int wn, int lr,
                                           int lk, f4 (&acc)[2][2]) {
#pragma unroll
  for (int kk = 0; kk < kDW; kk += 4) {
    float a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i] = sAt[kk + lk][wm + 16 * i + lr];
#pragma unroll
    for (int j = 0; j < 2; ++j) b[j] = sB[kk + lk][wn + 16 * j + lr];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
  }
}

// Step buffers of a diagonal-block factorisation (floats): [2][T][4096]
// copies of tile column k, the same for tile row k, [2][2][4096] inverses of
// tile k's factors (L, then U).
struct BlkBufs {
  float* col;
  float* row;
  float* inv;
  int T;
  float* Z;   // L11^-1 (npad x npad, the diagonal blocks written by the steps)
  float* lt;  // two-launch steps: L_ik and U_kj of step k ([T][4096] each)
  float* ut;
};

// Incremental inverses of a panel whose steps all take one launch: in step
// k's launch, tile column k of U11^-1 (Y) and tile row k of L11^-1 (Z), from
// tiles final in earlier launches:
//   Y[r, k] = -(sum_{t=r}^{k-1} Y[r, t] U[t, k]) Y[k, k]      (r < k)
//   Z[k, c] = -Z[k, k] (sum_{t=c}^{k-1} L[k, t] Z[t, c])      (c < k)
// with U[t, k], L[k, t] in W (solved in launch t) and Y[k, k], Z[k, k] the
// inverses of tile k's factors (its factorisation in launch k-1).  kY: Y's
// row tile o, else Z's column tile o.  s0, s1 are free LDS.
template <bool kY>
__device__ __forceinline__ void blk_inc_inverse(const float* __restrict__ W, long long ld, long long c0, int k, int o,
                                                float* __restrict__ Y, float* __restrict__ Z, int npad,
                                                float (*s0)[68], float (*s1)[68]) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const long long tk = 64LL * k, oo = 64LL * o;
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (f4){0, 0, 0, 0};
#pragma unroll 1
  for (int t = o; t < k; ++t) {
    const long long tt = 64LL * t;
    __syncthreads();  // the previous term's reads are done (before the loads:
                      // loaded values kept across a barrier are spilled)
    float4 va[4], vb[4];
    if constexpr (kY) {
      blk_load(Y + oo * npad + tt, npad, va);          // Y[r, t]
      blk_load(W + (c0 + tt) * ld + c0 + tk, ld, vb);  // U[t, k]
    } else {
      blk_load(W + (c0 + tk) * ld + c0 + tt, ld, va);  // L[k, t]
      blk_load(Z + tt * npad + oo, npad, vb);          // Z[t, c]
    }
    blk_put_t(s0, va);
    blk_put(s1, vb);
    __syncthreads();
    blk_mm_acc(s0, s1, wm, wn, lr, lk, acc);
  }
  // the sum times the diagonal inverse: Y: S Y[k, k]; Z: Z[k, k] S
  __syncthreads();
  float4 vd[4];
  blk_load((kY ? Y : Z) + tk * npad + tk, npad, vd);
  if constexpr (kY) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) s0[wn + 16 * j + lr][wm + 16 * i + 4 * lk + r] = acc[i][j][r];  // S^T
    blk_put(s1, vd);
  } else {
    blk_put_t(s0, vd);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) s1[wm + 16 * i + 4 * lk + r][wn + 16 * j + lr] = acc[i][j][r];  // S
  }
  __syncthreads();
  f4 res[2][2];
  blk_mm(s0, s1, wm, wn, lr, lk, res);
  float* out = kY ? Y + oo * npad + tk : Z + tk * npad + oo;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        out[(long long)(wm + 16 * i + 4 * lk + r) * npad + wn + 16 * j + lr] = -res[i][j][r];
}

// Step k of the diagonal block (see above); k = -1: workgroup 0 factors tile
// 0, workgroups 1 .. T-1 copy tile column 0 and T .. 2T-2 tile row 0 into the
// step buffers.  nt = T - 1 - k trailing tiles per side, workgroup b takes
// tile (k+1 + b / nt, k+1 + b % nt).
__global__ __launch_bounds__(256, 1) void k_blk_step(float* __restrict__ W, long long ld, long long c0, int k, int nt,
                                                  BlkBufs bb, float* __restrict__ Y, int npad, float* __restrict__ dg,
                                                  float* __restrict__ linv, int* __restrict__ ipiv,
                                                  unsigned* __restrict__ growth) {
  __shared__ __attribute__((aligned(16))) float s0[kDW][68];
  __shared__ __attribute__((aligned(16))) float s1[kDW][68];
  __shared__ __attribute__((aligned(16))) float s2[kDW][68];
  __shared__ __attribute__((aligned(16))) float s3[kDW][68];
  const int tid = threadIdx.x;
  constexpr int TS = kDW * kDW;
  if (k < 0) {
    const int T = bb.T;
    if (blockIdx.x == 0) {
      float4 v[4];
      blk_load(W + c0 * ld + c0, ld, v);
      blk_put(s0, v);
      __syncthreads();
      blk_factor_tile(s0, W, ld, c0, 0, bb.inv, Y, bb.Z, npad, dg, linv, ipiv, growth);
      return;
    }
    // tile (i, 0) or (0, j): row segments of 16 floats, one float4 per thread
    // and pass
    const bool colt = (int)blockIdx.x < T;
    const int t = colt ? (int)blockIdx.x : (int)blockIdx.x - T + 1;
    const float* src = colt ? W + (c0 + 64LL * t) * ld + c0 : W + c0 * ld + c0 + 64LL * t;
    float4* dst = reinterpret_cast<float4*>((colt ? bb.col : bb.row) + (size_t)t * TS);
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int idx = tid + 256 * p, r = idx / 16, c4 = (idx % 16) * 4;
      dst[idx] = *reinterpret_cast<const float4*>(src + r * ld + c4);
    }
    return;
  }
  if ((int)blockIdx.x >= nt * nt) {
    // incremental inverses (panels whose steps all take one launch)
    const int q = (int)blockIdx.x - nt * nt;
    if (q < k) blk_inc_inverse<true>(W, ld, c0, k, q, Y, bb.Z, npad, s0, s1);
    else blk_inc_inverse<false>(W, ld, c0, k, q - k, Y, bb.Z, npad, s0, s1);
    return;
  }
  const int par = k & 1, T = bb.T;
  const int ti = k + 1 + (int)blockIdx.x / nt, tj = k + 1 + (int)blockIdx.x % nt;
  const float* cur_inv = bb.inv + (size_t)par * 2 * TS;
  const float* cur_col = bb.col + ((size_t)par * T + ti) * TS;
  const float* cur_row = bb.row + ((size_t)par * T + tj) * TS;
  // every load in flight before the LDS stores
  float4 va[4], vu[4], vl[4], vb[4];
  blk_load(cur_col, kDW, va);        // A_ik
  blk_load(cur_inv + TS, kDW, vu);   // U_kk^-1
  blk_load(cur_inv, kDW, vl);        // L_kk^-1
  blk_load(cur_row, kDW, vb);        // A_kj
  const int wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const long long rC = c0 + 64LL * ti, cC = c0 + 64LL * tj;
  float cv[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) cv[i][j][r] = W[(rC + wm + 16 * i + 4 * lk + r) * ld + cC + wn + 16 * j + lr];
  blk_put_t(s0, va);
  blk_put(s1, vu);
  blk_put_t(s2, vl);
  blk_put(s3, vb);
  __syncthreads();
  f4 accL[2][2], accU[2][2];
  blk_mm(s0, s1, wm, wn, lr, lk, accL);  // L_ik = A_ik U_kk^-1
  blk_mm(s2, s3, wm, wn, lr, lk, accU);  // U_kj = L_kk^-1 A_kj
  __syncthreads();
  // L_ik as the next product's transposed left operand, U_kj as its right one
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        s0[wn + 16 * j + lr][wm + 16 * i + 4 * lk + r] = accL[i][j][r];
        s1[wm + 16 * i + 4 * lk + r][wn + 16 * j + lr] = accU[i][j][r];
      }
  if (tj == k + 1) {
    // tile (ti, k) of W takes L_ik
    float g = 0.0f;
    const long long rL = c0 + 64LL * ti, cL = c0 + 64LL * k;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          W[(rL + wm + 16 * i + 4 * lk + r) * ld + cL + wn + 16 * j + lr] = accL[i][j][r];
          g = abs_max_keep_nan(g, accL[i][j][r]);
        }
    ts_growth_max(g, growth);
  }
  if (ti == k + 1) {
    // tile (k, tj) of W takes U_kj
    const long long rU = c0 + 64LL * k, cU = c0 + 64LL * tj;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) W[(rU + wm + 16 * i + 4 * lk + r) * ld + cU + wn + 16 * j + lr] = accU[i][j][r];
  }
  __syncthreads();
  f4 acc[2][2];
  blk_mm(s0, s1, wm, wn, lr, lk, acc);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) cv[i][j][r] -= acc[i][j][r];
  if (blockIdx.x == 0) {
    // tile (k+1, k+1): factor the result
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) s2[wm + 16 * i + 4 * lk + r][wn + 16 * j + lr] = cv[i][j][r];
    __syncthreads();
    blk_factor_tile(s2, W, ld, c0, k + 1, bb.inv + (size_t)(par ^ 1) * 2 * TS, Y, bb.Z, npad, dg, linv, ipiv,
                    growth);
    return;
  }
  float* ncol = (tj == k + 1) ? bb.col + ((size_t)(par ^ 1) * T + ti) * TS : nullptr;
  float* nrow = (ti == k + 1) ? bb.row + ((size_t)(par ^ 1) * T + tj) * TS : nullptr;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rr = wm + 16 * i + 4 * lk + r, cc = wn + 16 * j + lr;
        W[(rC + rr) * ld + cC + cc] = cv[i][j][r];
        if (ncol) ncol[rr * kDW + cc] = cv[i][j][r];
        if (nrow) nrow[rr * kDW + cc] = cv[i][j][r];
      }
}

// Two-launch form of step k, for steps with many trailing tiles (the fused
// step forms each L_ik and U_kj again in every workgroup of its tile row or
// column: three products per trailing tile).  k_blk_solve2 forms tile column
// and row k once (workgroups [0, nt): L_ik into W and bb.lt, with their
// growth; [nt, 2 nt): U_kj into W and bb.ut); k_blk_update2 then subtracts
// one product per trailing tile, and its workgroup 0 factors tile (k+1, k+1)
// as the fused step's does.
__global__ __launch_bounds__(256) void k_blk_solve2(float* __restrict__ W, long long ld, long long c0, int k, int nt,
                                                    BlkBufs bb, unsigned* __restrict__ growth) {
  __shared__ __attribute__((aligned(16))) float s0[kDW][68];
  __shared__ __attribute__((aligned(16))) float s1[kDW][68];
  constexpr int TS = kDW * kDW;
  const int par = k & 1, T = bb.T;
  const bool lcol = (int)blockIdx.x < nt;
  const int q = lcol ? (int)blockIdx.x : (int)blockIdx.x - nt;  // tile k+1+q of the column / row
  const float* inv = bb.inv + (size_t)par * 2 * TS;
  float4 va[4], vb[4];
  if (lcol) {
    blk_load(bb.col + ((size_t)par * T + k + 1 + q) * TS, kDW, va);  // A_ik
    blk_load(inv + TS, kDW, vb);                                     // U_kk^-1
  } else {
    blk_load(inv, kDW, va);                                          // L_kk^-1
    blk_load(bb.row + ((size_t)par * T + k + 1 + q) * TS, kDW, vb);  // A_kj
  }
  blk_put_t(s0, va);
  blk_put(s1, vb);
  __syncthreads();
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f4 acc[2][2];
  blk_mm(s0, s1, wm, wn, lr, lk, acc);
  const long long r0 = c0 + 64LL * (lcol ? k + 1 + q : k), cc0 = c0 + 64LL * (lcol ? k : k + 1 + q);
  float* out = (lcol ? bb.lt : bb.ut) + (size_t)q * TS;
  float g = 0.0f;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rr = wm + 16 * i + 4 * lk + r, cc = wn + 16 * j + lr;
        W[(r0 + rr) * ld + cc0 + cc] = acc[i][j][r];
        out[rr * kDW + cc] = acc[i][j][r];
        g = abs_max_keep_nan(g, acc[i][j][r]);
      }
  if (lcol) ts_growth_max(g, growth);
}

__global__ __launch_bounds__(256, 1) void k_blk_update2(float* __restrict__ W, long long ld, long long c0, int k,
                                                        int nt, BlkBufs bb, float* __restrict__ Y, int npad,
                                                        float* __restrict__ dg, float* __restrict__ linv,
                                                        int* __restrict__ ipiv, unsigned* __restrict__ growth) {
  __shared__ __attribute__((aligned(16))) float s0[kDW][68];
  __shared__ __attribute__((aligned(16))) float s1[kDW][68];
  __shared__ __attribute__((aligned(16))) float s2[kDW][68];
  constexpr int TS = kDW * kDW;
  const int par = k & 1, T = bb.T;
  const int i = (int)blockIdx.x / nt, j = (int)blockIdx.x % nt;
  const int ti = k + 1 + i, tj = k + 1 + j;
  float4 va[4], vb[4];
  blk_load(bb.lt + (size_t)i * TS, kDW, va);
  blk_load(bb.ut + (size_t)j * TS, kDW, vb);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const long long rC = c0 + 64LL * ti, cC = c0 + 64LL * tj;
  float cv[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) cv[a][b][r] = W[(rC + wm + 16 * a + 4 * lk + r) * ld + cC + wn + 16 * b + lr];
  blk_put_t(s0, va);
  blk_put(s1, vb);
  __syncthreads();
  f4 acc[2][2];
  blk_mm(s0, s1, wm, wn, lr, lk, acc);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) cv[a][b][r] -= acc[a][b][r];
  if (blockIdx.x == 0) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) s2[wm + 16 * a + 4 * lk + r][wn + 16 * b + lr] = cv[a][b][r];
    __syncthreads();
    blk_factor_tile(s2, W, ld, c0, k + 1, bb.inv + (size_t)(par ^ 1) * 2 * TS, Y, bb.Z, npad, dg, linv, ipiv,
                    growth);
    return;
  }
  float* ncol = (tj == k + 1) ? bb.col + ((size_t)(par ^ 1) * T + ti) * TS : nullptr;
  float* nrow = (ti == k + 1) ? bb.row + ((size_t)(par ^ 1) * T + tj) * TS : nullptr;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rr = wm + 16 * a + 4 * lk + r, cc = wn + 16 * b + lr;
        W[(rC + rr) * ld + cC + cc] = cv[a][b][r];
        if (ncol) ncol[rr * kDW + cc] = cv[a][b][r];
        if (nrow) nrow[rr * kDW + cc] = cv[a][b][r];
      }
}

// Identity in the diagonal blocks [t0, npad / 64) of Y and Z (the padding
// of U11^-1 and L11^-1).
__global__ __launch_bounds__(256) void k_binv_identity(int t0, int npad, float* __restrict__ Y, float* __restrict__ Z) {
  const long long b0 = 64LL * (t0 + blockIdx.x);
  for (int i = threadIdx.x; i < kDW * kDW; i += 256) {
    const long long o = (b0 + i / kDW) * npad + b0 + i % kDW;
    Y[o] = Z[o] = (i / kDW == i % kDW);
  }
}

// One batched FP32 product C_p = alpha * A_p B_p (C not read), each operand
// of batch p at base + p * stride; M, N, K multiples of the tile.  A's
// entries at row ar0 + p * astep + r >= alim or column ac0 + p * astep + k >=
// alim read as zero (the identity padding of the factors' off-diagonal
// blocks).  triA / triB: 1 = that operand is upper triangular, 2 = lower; its
// entries on the other side of the diagonal read as zero (they are not
// stored).
struct BatOp {
  const float* A;
  long long lda, sa;
  int ar0, ac0, astep, alim, triA;
  const float* B;
  long long ldb, sb;
  int triB;
  float* C;
  long long ldc, sc;
  float alpha;
};

__device__ __forceinline__ bool bat_keep(int tri, int r, int c) { return tri == 0 || (tri == 1 ? c >= r : c <= r); }

// Two batched products of the same shape in one launch: blockIdx.z < P takes
// o0's batch blockIdx.z, the others o1's.  TM x TM tile per 256-thread
// workgroup, 2 x 2 waves of (TM/2)^2, MFMA f32 16x16x4, K staged through LDS
// in steps of 32, the next step's loads in flight during this one's MFMAs.
template <int TM>
__global__ __launch_bounds__(256) void k_gemm_bat(int K, int P, BatOp o0, BatOp o1) {
  constexpr int BK = 32, MI = TM / 32, PAD = 4;
  __shared__ float sA[BK][TM + PAD];
  __shared__ float sB[BK][TM + PAD];
  const bool second = (int)blockIdx.z >= P;
  const BatOp& o = second ? o1 : o0;
  const int p = second ? (int)blockIdx.z - P : (int)blockIdx.z;
  const float* A = o.A + (long long)p * o.sa;
  const float* B = o.B + (long long)p * o.sb;
  float* C = o.C + (long long)p * o.sc;
  const int rlim = o.alim - (o.ar0 + p * o.astep), klim = o.alim - (o.ac0 + p * o.astep);
  const int triA = o.triA, triB = o.triB;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TM;
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int wm = (wave / 2) * (TM / 2), wn = (wave % 2) * (TM / 2);
  const int lr = lane % 16, lk = lane / 16;
  f4 acc[MI][MI];
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int j = 0; j < MI; ++j) acc[i][j] = (f4){0, 0, 0, 0};
  constexpr int NA = TM * BK / 256;  // elements per thread and operand
  // the next K step's operands are loaded into registers while this one's
  // MFMAs run
  float va[NA], vb[NA];
  auto load = [&](int k0) {
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      const int idx = tid + 256 * i;
      const int r = m0 + idx / BK, k = k0 + idx % BK;
      va[i] = (r < rlim && k < klim && bat_keep(triA, r, k)) ? A[(long long)r * o.lda + k] : 0.0f;
    }
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      const int idx = tid + 256 * i;
      const int k = k0 + idx / TM, n = n0 + idx % TM;
      vb[i] = bat_keep(triB, k, n) ? B[(long long)k * o.ldb + n] : 0.0f;
    }
  };
  // a triangular operand's zero triangle is skipped: K runs only over the
  // steps where the tile's rows of A / columns of B can be nonzero
  int kb = 0, ke = K;
  if (triA == 1) kb = max(kb, m0);       // A upper: A[r][k] = 0 for k < r
  if (triA == 2) ke = min(ke, m0 + TM);  // A lower: A[r][k] = 0 for k > r
  if (triB == 1) ke = min(ke, n0 + TM);  // B upper: B[k][n] = 0 for n < k
  if (triB == 2) kb = max(kb, n0);       // B lower: B[k][n] = 0 for n > k
  kb = kb / BK * BK;
  if (kb < ke) load(kb);
  for (int k0 = kb; k0 < ke; k0 += BK) {
    __syncthreads();  // the previous step's fragment reads are done
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      const int idx = tid + 256 * i;
      sA[idx % BK][idx / BK] = va[i];
      sB[idx / TM][idx % TM] = vb[i];
    }
    __syncthreads();
    if (k0 + BK < ke) load(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      float a[MI], b[MI];
#pragma unroll
      for (int i = 0; i < MI; ++i) a[i] = sA[kk + lk][wm + 16 * i + lr];
#pragma unroll
      for (int j = 0; j < MI; ++j) b[j] = sB[kk + lk][wn + 16 * j + lr];
#pragma unroll
      for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < MI; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int j = 0; j < MI; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        C[(long long)(m0 + wm + 16 * i + 4 * lk + r) * o.ldc + n0 + wn + 16 * j + lr] = o.alpha * acc[i][j][r];
}

// A16[i][k] = -Z[i][k] for k < i < jb (Z = L11^-1, row pitch npad), zero
// elsewhere up to Kp: the strictly lower part of L11^-1, negated, so that
// C -= A16 B16^T on C = A12 leaves L11^-1 A12.  One thread per 8 halves.
__global__ __launch_bounds__(256) void k_linv_a16(int jb, int npad, int Kp, const float* __restrict__ Z,
                                                  _Float16* __restrict__ A16) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const int per = Kp / 8;
  if (idx >= (long long)jb * per) return;
  const int i = (int)(idx / per), k0 = (int)(idx % per) * 8;
  h8 v;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int k = k0 + q;
    v[q] = (_Float16)((k < i && k < jb) ? -Z[(long long)i * npad + k] : 0.0f);
  }
  *(h8*)(A16 + (long long)i * Kp + k0) = v;
}

// Y[k][n] = dg[k] Y[k][n] for k < n < jb, 0 for n <= k (row pitch npad, a
// multiple of 64): U11^-1 becomes the inverse of U11 with its columns divided
// by their diagonal entries (unit upper triangular), less the identity.
__global__ __launch_bounds__(256) void k_binv_rowscale(int jb, int npad, float* __restrict__ Y,
                                                       const float* __restrict__ dg) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const int per = npad / 4;
  if (idx >= (long long)jb * per) return;
  const int k = (int)(idx / per), n = (int)(idx % per) * 4;
  if (n >= jb) return;
  float4* p = reinterpret_cast<float4*>(Y + (long long)k * npad + n);
  const float d = dg[k];
  float4 v = *p;
  v.x = (n > k) ? v.x * d : 0.0f;
  v.y = (n + 1 > k) ? v.y * d : 0.0f;
  v.z = (n + 2 > k) ? v.z * d : 0.0f;
  v.w = (n + 3 > k) ? v.w * d : 0.0f;
  *p = v;
}

// A16[r][k] = -A[r][k] / u_kk (nrd[k] = -1 / u_kk) for k < K, zero up to Kp,
// and A[r][k] itself replaced by A[r][k] / u_kk (FP32): with N the strictly
// upper part of the unit upper triangular inverse of U11 diag(1/u_kk),
// L21 = A21 U11^-1 = Ahat + Ahat N for Ahat = A21 diag(1/u_kk), and only the
// correction Ahat N goes through FP16 (C -= colscale * (A16 B16^T)).  One
// thread per 8 halves; vec: A and lda 16-byte aligned and K a multiple of 8.
__global__ __launch_bounds__(256) void k_cvt_a16_l21(long long M, int K, int Kp, float* __restrict__ A, long long lda,
                                                     const float* __restrict__ nrd, _Float16* __restrict__ A16,
                                                     int vec) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const int per = Kp / 8;
  if (idx >= M * per) return;
  const long long r = idx / per;
  const int k0 = (int)(idx % per) * 8;
  h8 v;
  if (vec && k0 + 8 <= K) {
    float* src = A + r * lda + k0;
    const float4 p0 = *reinterpret_cast<const float4*>(src);
    const float4 p1 = *reinterpret_cast<const float4*>(src + 4);
    const float4 s0 = *reinterpret_cast<const float4*>(nrd + k0);
    const float4 s1 = *reinterpret_cast<const float4*>(nrd + k0 + 4);
    v[0] = (_Float16)(p0.x * s0.x);
    v[1] = (_Float16)(p0.y * s0.y);
    v[2] = (_Float16)(p0.z * s0.z);
    v[3] = (_Float16)(p0.w * s0.w);
    v[4] = (_Float16)(p1.x * s1.x);
    v[5] = (_Float16)(p1.y * s1.y);
    v[6] = (_Float16)(p1.z * s1.z);
    v[7] = (_Float16)(p1.w * s1.w);
    *reinterpret_cast<float4*>(src) = make_float4(-p0.x * s0.x, -p0.y * s0.y, -p0.z * s0.z, -p0.w * s0.w);
    *reinterpret_cast<float4*>(src + 4) = make_float4(-p1.x * s1.x, -p1.y * s1.y, -p1.z * s1.z, -p1.w * s1.w);
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k = k0 + q;
      float x = 0.0f;
      if (k < K) {
        x = A[r * lda + k] * nrd[k];
        A[r * lda + k] = -x;
      }
      v[q] = (_Float16)x;
    }
  }
  *(h8*)(A16 + r * Kp + k0) = v;
}

// k_cvt_a16 that also folds the largest |a| (NaN kept) into *growth: the
// trailing update's conversion of L21, which reads every multiplier below a
// diagonal-pivot panel once.
__global__ __launch_bounds__(256) void k_cvt_a16_growth(long long M, int K, int Kp, const float* __restrict__ A,
                                                        long long lda, _Float16* __restrict__ A16, int vec,
                                                        unsigned* __restrict__ growth) {
  __shared__ unsigned wmax[4];
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const int per = Kp / 8;
  float g = 0.0f;
  if (idx < M * per) {
    const long long r = idx / per;
    const int k0 = (int)(idx % per) * 8;
    float x[8];
    if (vec && k0 + 8 <= K) {
      const float4 p0 = *reinterpret_cast<const float4*>(A + r * lda + k0);
      const float4 p1 = *reinterpret_cast<const float4*>(A + r * lda + k0 + 4);
      x[0] = p0.x;
      x[1] = p0.y;
      x[2] = p0.z;
      x[3] = p0.w;
      x[4] = p1.x;
      x[5] = p1.y;
      x[6] = p1.z;
      x[7] = p1.w;
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) x[q] = (k0 + q < K) ? A[r * lda + k0 + q] : 0.0f;
    }
    h8 v;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      v[q] = (_Float16)x[q];
      g = abs_max_keep_nan(g, x[q]);
    }
    *(h8*)(A16 + r * Kp + k0) = v;
  }
  ts_growth_max_wg(g, growth, wmax);
}

// ---------------------------------------------------------------------------
// host-side LU driver
template <typename T>
struct LuCtx {
  qd_context* h;
  qd_lu_state* s;
  T* W;
  long long ld, N;
  bool f16;   // trailing updates in fp16
  bool calu;  // tournament pivoting in the leaves (preconditioner factorisations)
  hipStream_t st = nullptr;                   // stream of every launch of this context
  qd_lu_state::F16Scratch* sc = nullptr;      // FP16 operand copies of this context
  qd_lu_state::F16Scratch* sc_solve = nullptr;  // those of trsm_lower's inner updates (nullptr: sc)
  bool diag = false;  // with calu: the leaves take their diagonal pivots instead (checked by growth)
  const double* src = nullptr;  // column-major FP64 A that the factorisation first converts into W
  long long src_ld = 0;
  // the current step's L21 in FP16 ([rows][Kp], from its panel's product):
  // trailing updates read it instead of converting L21 (nullptr: convert)
  const _Float16* a16 = nullptr;
};

// The FP16 update C -= (A16 B16^T) * colscale on stream st; max_wgs > 0
// caps the workgroups (each then loops over several tiles).
// K of the FP16 operand copies, zero-padded to a multiple of 64 (the wide-tile
// kernel's K step; k_gemm16 takes any multiple of 32).
static inline int f16_kpad(int K) { return (K + 63) / 64 * 64; }

// Launches with at least this many 256 x 256 tiles take the wide-tile kernel
// (QUADESTH_GEMM16W_MIN; 0 turns it off).  Measured on the 76,806-DOF
// factorisation's update shapes: 7% faster than k_gemm16 on the square
// updates, 2% slower on the 2048-column ones (9 waves of tiles), slower on
// small launches.
static long long wide_gemm_min_tiles() {
  static const long long v = [] {
    const char* e = std::getenv("QUADESTH_GEMM16W_MIN");
    return e ? std::atoll(e) : 1024LL;
  }();
  return v;
}

// o16.h non-null: the result also goes to o16 as FP16, its largest |c| to
// o16.growth (kernel role 2).
// ktri (tile_kext): a triangular operand's zero part is skipped per tile.
static int launch_f16_update(qd_context* h, bool trailing, long long M, long long N, int Kp, const _Float16* A16,
                             const _Float16* B16, const float* colscale, float* C, long long ldc, int rows16,
                             int max_wgs, hipStream_t st, Out16 o16 = Out16{nullptr, 0, nullptr}, int ktri = 0) {
  const long long wmin = wide_gemm_min_tiles();
  const long long wtiles = ((N + w16::BN - 1) / w16::BN) * ((M + w16::BM - 1) / w16::BM);
  // the wide kernel takes K steps of 64 and 32-bit element offsets into A16 / B16
  if (wmin > 0 && wtiles >= wmin && max_wgs <= 0 && Kp % w16::BK == 0 &&
      std::max(M, N) * (long long)Kp < (1LL << 32)) {
    if (o16.h)
      hipLaunchKernelGGL(k_gemm16w<2>, dim3((unsigned)wtiles), dim3(256), 0, st, M, N, Kp, A16, B16, colscale, C, ldc,
                         rows16, o16, ktri);
    else if (trailing)
      hipLaunchKernelGGL(k_gemm16w<1>, dim3((unsigned)wtiles), dim3(256), 0, st, M, N, Kp, A16, B16, colscale, C, ldc,
                         rows16, o16, ktri);
    else
      hipLaunchKernelGGL(k_gemm16w<0>, dim3((unsigned)wtiles), dim3(256), 0, st, M, N, Kp, A16, B16, colscale, C, ldc,
                         rows16, o16, ktri);
    QD_HIP(h, hipGetLastError());
    return QD_OK;
  }
  const long long tiles = ((N + H_BN - 1) / H_BN) * ((M + H_BM - 1) / H_BM);
  const dim3 g((unsigned)(max_wgs > 0 ? std::min<long long>(tiles, max_wgs) : tiles));
  if (o16.h)
    hipLaunchKernelGGL(k_gemm16<2>, g, dim3(256), 0, st, M, N, Kp, A16, B16, colscale, C, ldc, rows16, o16, ktri);
  else if (trailing)
    hipLaunchKernelGGL(k_gemm16<1>, g, dim3(256), 0, st, M, N, Kp, A16, B16, colscale, C, ldc, rows16, o16, ktri);
  else
    hipLaunchKernelGGL(k_gemm16<0>, g, dim3(256), 0, st, M, N, Kp, A16, B16, colscale, C, ldc, rows16, o16, ktri);
  QD_HIP(h, hipGetLastError());
  return QD_OK;
}

// QUADESTH_NARROW_GEMM=0 sends the narrow FP32 updates to the square-tile kernel.
static bool narrow_gemm() {
  static const bool on = [] {
    const char* e = std::getenv("QUADESTH_NARROW_GEMM");
    return !(e && e[0] == '0');
  }();
  return on;
}

// Non-trailing FP32 updates of at least this many flops go through FP16
// operand copies (QUADESTH_F16_MIN_WORK overrides it).
static double f16_min_work() {
  static const double v = [] {
    const char* e = std::getenv("QUADESTH_F16_MIN_WORK");
    return e ? std::atof(e) : 2.0e9;
  }();
  return v;
}
// QUADESTH_F16_INNER: FP16 for the inner updates of the triangular solves
// (bit 0) and of the panels (bit 1); default both.
static int f16_inner() {
  static const int v = [] {
    const char* e = std::getenv("QUADESTH_F16_INNER");
    return e ? std::atoi(e) : 3;
  }();
  return v;
}
constexpr int kStripMaxRows = 1024;      // LDS of k_trsm_strip<8>: 32 KiB
constexpr long long kStripNarrowCols = 8192;

// reuse_a16: A's FP16 copy is already in c.sc->h16a (the same A as the
// previous call on this context's stream).
// max_wgs > 0 caps the workgroups of the FP16 kernel (each then loops over
// several tiles).
// growth: the largest |a| of A goes there as well (FP16 path only).
template <typename T>
static int gemm_sub(LuCtx<T>& c, long long M, long long Nn, int K, const T* A, const T* B, T* C, bool allow_f16,
                    bool trailing = false, bool reuse_a16 = false, int max_wgs = 0, unsigned* growth = nullptr) {
  if (M <= 0 || Nn <= 0 || K <= 0) return QD_OK;
  int rc_ = QD_OK;
  dim3 grid((unsigned)((Nn + GBN - 1) / GBN), (unsigned)((M + GBM - 1) / GBM));
  if constexpr (std::is_same<T, float>::value) {
    // FP16 conversion costs three extra launches; below f16_min_work() flops
    // the FP32 MFMA kernel on W directly is faster.
    if (c.f16 && allow_f16 && (trailing || (double)M * (double)Nn * (double)K >= f16_min_work())) {
      const long long nb = (Nn + 63) / 64;
      const int Kp = f16_kpad(K);
      QD_ALLOC(c.h, c.sc->colscale, (size_t)nb);
      QD_ALLOC(c.h, c.sc->h16a, (size_t)M * Kp);
      QD_ALLOC(c.h, c.sc->h16b, (size_t)Nn * Kp);
      const int vb = ((reinterpret_cast<uintptr_t>(B) & 15) == 0 && c.ld % 4 == 0) ? 1 : 0;
      const int va = ((reinterpret_cast<uintptr_t>(A) & 15) == 0 && c.ld % 4 == 0 && K % 8 == 0) ? 1 : 0;
      hipLaunchKernelGGL(k_scale_cvt_b16t, dim3((unsigned)nb), dim3(256), 0, c.st, K, Nn, Kp, B, c.ld,
                         c.sc->colscale.p, c.sc->h16b.p, vb);
      // a trailing update with its panel's FP16 L21 at hand converts nothing
      const _Float16* a16 = (trailing && c.a16) ? c.a16 : c.sc->h16a.p;
      if (trailing && c.a16) reuse_a16 = true;
      if (!reuse_a16 && growth)
        hipLaunchKernelGGL(k_cvt_a16_growth, dim3((unsigned)((M * (Kp / 8) + 255) / 256)), dim3(256), 0, c.st, M,
                           K, Kp, A, c.ld, c.sc->h16a.p, va, growth);
      else if (!reuse_a16)
        hipLaunchKernelGGL(k_cvt_a16, dim3((unsigned)((M * (Kp / 8) + 255) / 256)), dim3(256), 0, c.st, M, K,
                           Kp, A, c.ld, c.sc->h16a.p, va);
      ProfScope ps(c.h, trailing ? QD_PROF_LU_TRAILING_GEMM : -1, 2.0 * (double)M * (double)Nn * K, c.st);
      const int rows16 = (c.ld % 4 == 0 && (reinterpret_cast<uintptr_t>(C) & 15) == 0) ? 1 : 0;
      if ((rc_ = launch_f16_update(c.h, trailing, M, Nn, Kp, a16, c.sc->h16b.p, c.sc->colscale.p, C, c.ld, rows16,
                                   max_wgs, c.st)))
        return rc_;
      QD_HIP(c.h, hipGetLastError());
      return QD_OK;
    }
  }
  if constexpr (std::is_same<T, double>::value) {
    dim3 g64((unsigned)((Nn + D_BN - 1) / D_BN), (unsigned)((M + D_BM - 1) / D_BM));
    ProfScope ps(c.h, trailing ? QD_PROF_LU_TRAILING_GEMM : -1, 2.0 * (double)M * (double)Nn * K, c.st);
    if (trailing)
      hipLaunchKernelGGL(k_gemm64<1>, g64, dim3(256), 0, c.st, M, Nn, K, A, c.ld, B, c.ld, C, c.ld);
    else
      hipLaunchKernelGGL(k_gemm64<0>, g64, dim3(256), 0, c.st, M, Nn, K, A, c.ld, B, c.ld, C, c.ld);
    QD_HIP(c.h, hipGetLastError());
    return QD_OK;
  }
  ProfScope ps(c.h, trailing ? QD_PROF_LU_TRAILING_GEMM : -1, 2.0 * (double)M * (double)Nn * K, c.st);
  if constexpr (std::is_same<T, float>::value) {
    // narrow updates (inside the panels): one launch of the column-narrow kernel
    if (Nn <= 64 && narrow_gemm()) {
      const dim3 gn((unsigned)((M + 127) / 128));
      if (Nn <= 32)
        hipLaunchKernelGGL(k_gemm_narrow<32>, gn, dim3(256), 0, c.st, M, (int)Nn, K, A, c.ld, B, c.ld, C, c.ld);
      else
        hipLaunchKernelGGL(k_gemm_narrow<64>, gn, dim3(256), 0, c.st, M, (int)Nn, K, A, c.ld, B, c.ld, C, c.ld);
      QD_HIP(c.h, hipGetLastError());
      return QD_OK;
    }
  }
  hipLaunchKernelGGL(k_gemm_sub<T>, grid, dim3(256), 0, c.st, M, Nn, K, A, c.ld, B, c.ld, C, c.ld);
  QD_HIP(c.h, hipGetLastError());
  return QD_OK;
}

// X = L^-1 B for a diagonal-block panel's L11 (rows [r0, r0+m)) whose
// inverse's strictly lower part, negated, is at li in FP16 ([m][Kp]): X = B +
// (L11^-1 - I) B as one FP16 update of B in place (only the correction goes
// through FP16).  The B operand copy goes to the solve's scratch, sized
// before the factorisation.
static int trsm_lower_inv(LuCtx<float>& c, long long r0, int m, long long c0, long long c1, const _Float16* li) {
  qd_lu_state::F16Scratch* sc = c.sc_solve ? c.sc_solve : c.sc;
  const long long n = c1 - c0, nb = (n + 63) / 64;
  const int Kp = f16_kpad(m);
  if (sc->h16b.count < (size_t)(n * Kp) || sc->colscale.count < (size_t)nb)
    return set_error(c.h, QD_ERR_STATE, "U12 solve scratch sized for fewer than %lld x %d", n, Kp);
  float* B = c.W + r0 * c.ld + c0;
  const int vb = ((reinterpret_cast<uintptr_t>(B) & 15) == 0 && c.ld % 4 == 0) ? 1 : 0;
  hipLaunchKernelGGL(k_scale_cvt_b16t, dim3((unsigned)nb), dim3(256), 0, c.st, m, n, Kp, (const float*)B, c.ld,
                     sc->colscale.p, sc->h16b.p, vb);
  QD_HIP(c.h, hipGetLastError());
  const int rows16 = (c.ld % 4 == 0 && (reinterpret_cast<uintptr_t>(B) & 15) == 0) ? 1 : 0;
  // L11^-1 - I is strictly lower triangular: a tile's K runs below its last row
  return launch_f16_update(c.h, false, m, n, Kp, li, sc->h16b.p, sc->colscale.p, B, c.ld, rows16, 0, c.st,
                           Out16{nullptr, 0, nullptr}, 2);
}

// QUADESTH_INV_U12=0 keeps the strip solves for U12 where a panel's L11^-1
// is at hand.
static bool inv_u12() {
  const char* e = std::getenv("QUADESTH_INV_U12");
  return !(e && e[0] == '0');
}

// X = L^-1 B, L unit lower at W[r0:r0+m, r0:r0+m], B = W[r0:r0+m, c0:c1]
template <typename T>
static int trsm_lower(LuCtx<T>& c, long long r0, int m, long long c0, long long c1) {
  if (c1 <= c0 || m <= 0) return QD_OK;
  if constexpr (std::is_same<T, float>::value) {
    if (c.f16 && inv_u12())
      for (int b = 0; b < 2; ++b)
        if (c.s->linvh_col[b] == r0 && c.s->linvh_w[b] == m) return trsm_lower_inv(c, r0, m, c0, c1, c.s->linvh[b].p);
    // One strip solve when there are few columns (narrow strips: more
    // workgroups); for many columns the recursion below stops at 256 rows,
    // where wide strips re-read little of L, and the GEMMs between carry the
    // bulk of the work.
    const long long n = c1 - c0;
    if (c.calu && m % 32 == 0 && r0 % 32 == 0 && (n <= kStripNarrowCols ? m <= kStripMaxRows : m <= 256)) {
      if (n <= kStripNarrowCols) {
        hipLaunchKernelGGL(k_trsm_strip<8>, dim3((unsigned)((n + 7) / 8)), dim3(256), (size_t)m * 8 * sizeof(float),
                           c.st, c.W, c.ld, r0, m, c0, c1, (const float*)c.s->ts_linv.p);
      } else {
        hipLaunchKernelGGL(k_trsm_strip<32>, dim3((unsigned)((n + 31) / 32)), dim3(256),
                           (size_t)m * 32 * sizeof(float), c.st, c.W, c.ld, r0, m, c0, c1,
                           (const float*)c.s->ts_linv.p);
      }
      QD_HIP(c.h, hipGetLastError());
      return QD_OK;
    }
  }
  if (m <= 32) {
    hipLaunchKernelGGL(k_trsm_small<T>, dim3((unsigned)((c1 - c0 + 255) / 256)), dim3(256), 0, c.st, c.W,
                       c.ld, r0, m, c0, c1);
    QD_HIP(c.h, hipGetLastError());
    return QD_OK;
  }
  int m1 = ((m / 2) + 31) / 32 * 32;
  int rc = trsm_lower(c, r0, m1, c0, c1);
  if (rc) return rc;
  // B2 -= L21 X1
  LuCtx<T> cs = c;
  if (c.sc_solve) cs.sc = c.sc_solve;
  rc = gemm_sub(cs, m - m1, c1 - c0, m1, c.W + (r0 + m1) * c.ld + r0, c.W + r0 * c.ld + c0,
                c.W + (r0 + m1) * c.ld + c0, (f16_inner() & 1) != 0);
  if (rc) return rc;
  return trsm_lower(c, r0 + m1, m - m1, c0, c1);
}

template <typename T>
static int colstep(LuCtx<T>& c, long long col, long long prev, long long lstart, long long lend, long long pstart,
                   long long pend) {
  const long long rows = c.N - col;
  if (rows <= 0) return QD_OK;
  const unsigned nb = (unsigned)((rows + kColRowsPerBlock - 1) / kColRowsPerBlock);
  hipLaunchKernelGGL(k_colstep<T>, dim3(nb), dim3(kColThreads), 0, c.st, c.W, c.ld, c.N, col, prev, lstart,
                     lend, pstart, pend, c.s->piv_part.p, c.s->piv_idx.p, c.s->piv_state.p, c.s->rpiv.p,
                     c.s->ipiv_dev.p);
  QD_HIP(c.h, hipGetLastError());
  return QD_OK;
}

// Leaf of <= 32 columns by tournament pivoting: selection levels, the
// interchange sequence, the interchanges on the panel columns, then the LU of
// the winners and the multipliers of all other rows.
static int tslu_leaf(LuCtx<float>& c, long long c0, int w, long long p0, long long p1) {
  qd_lu_state* s = c.s;
  const long long m = c.N - c0;
  long long ncand = 0, ncnt = 0;
  ts_tree_sizes(m, ncand, ncnt);
  if ((long long)s->ts_cand.count < ncand || (long long)s->ts_cnt.count < ncnt)
    return set_error(c.h, QD_ERR_STATE, "tournament buffers sized for fewer rows than %lld", m);
  using Nd = TsNode<kTsTreeR, kTsTreeV>;
  const unsigned nb = (unsigned)((m + Nd::rows - 1) / Nd::rows);
  if (c.diag) {
    // diagonal pivots (leaves of up to kDW columns): no interchanges
    if (w > kDW) return set_error(c.h, QD_ERR_STATE, "diagonal-pivot leaf of %d columns", w);
    const long long below = std::max<long long>(m - w, 1);
    float* linv = s->ts_linv.p + (c0 / kTsW) * kTsW * kTsW;
    if (w == kDW && m >= kDW) {
      hipLaunchKernelGGL(k_leaf_diag64, dim3(1), dim3(64), 0, c.st, (const float*)c.W, c.ld, c0, s->ts_lu.p, linv,
                         s->ipiv_dev.p, s->growth.p);
      hipLaunchKernelGGL(k_finish_diag64, dim3((unsigned)((below + 255) / 256)), dim3(256), 0, c.st, c.W, c.ld, c.N,
                         c0, (const float*)s->ts_lu.p, s->growth.p);
    } else {
      hipLaunchKernelGGL(k_leaf_diag, dim3(1), dim3(64), 0, c.st, (const float*)c.W, c.ld, c.N, c0, w, s->ts_lu.p, linv,
                         s->ipiv_dev.p, s->growth.p);
      hipLaunchKernelGGL(k_finish_diag, dim3((unsigned)((below + 255) / 256)), dim3(256), 0, c.st, c.W, c.ld, c.N,
                         c0, w, (const float*)s->ts_lu.p, s->growth.p);
    }
    QD_HIP(c.h, hipGetLastError());
    return QD_OK;
  }
  hipLaunchKernelGGL((k_tslu_tree<kTsTreeR, kTsTreeV>), dim3(nb), dim3(Nd::threads), 0, c.st, c.W, c.ld, c.N, c0, w, s->ts_cand.p, s->ts_cnt.p,
                     s->ts_lu.p, s->ts_top.p, s->ts_meta.p, s->ipiv_dev.p);
  const long long below = std::max<long long>(m - w, 1);
  const int row_blocks = (int)((below + 255) / 256);
  const int col_blocks = (int)((p1 - p0 + 255) / 256);
  hipLaunchKernelGGL(k_tslu_finish, dim3((unsigned)(row_blocks + col_blocks + 1)), dim3(256), 0, c.st, c.W, c.ld, c.N,
                     c0, w, p0, p1, (const float*)s->ts_lu.p, (const float*)s->ts_top.p, (const int*)s->ts_meta.p,
                     row_blocks, s->ts_linv.p + (c0 / kTsW) * kTsW * kTsW);
  QD_HIP(c.h, hipGetLastError());
  return QD_OK;
}

// Factor W[c0:N, c0:c0+w] (recursively), swaps over panel columns [p0, p1).
template <typename T>
static int panel(LuCtx<T>& c, long long c0, int w, long long p0, long long p1) {
  int rc;
  if constexpr (std::is_same<T, float>::value) {
    if (w <= (c.diag ? kDW : kLeaf) && c.calu) return tslu_leaf(c, c0, w, p0, p1);
  }
  if (w <= kLeaf) {
    const long long lend = c0 + w;
    for (long long col = c0; col <= lend; ++col) {
      if (col >= c.N && col - 1 >= c.N) break;
      rc = colstep(c, std::min(col, c.N), col > c0 ? col - 1 : -1, c0, std::min(lend, c.N), p0, p1);
      if (rc) return rc;
    }
    return QD_OK;
  }
  const int lw = (std::is_same<T, float>::value && c.calu && c.diag) ? kDW : kLeaf;
  const int w1 = ((w / 2) + lw - 1) / lw * lw;
  if ((rc = panel(c, c0, w1, p0, p1))) return rc;
  if ((rc = trsm_lower(c, c0, w1, c0 + w1, c0 + w))) return rc;
  if ((rc = gemm_sub(c, c.N - c0 - w1, w - w1, w1, c.W + (c0 + w1) * c.ld + c0, c.W + c0 * c.ld + c0 + w1,
                     c.W + (c0 + w1) * c.ld + c0 + w1, (f16_inner() & 2) != 0)))
    return rc;
  return panel(c, c0 + w1, w - w1, p0, p1);
}

// QUADESTH_BLOCK_PANEL=0 keeps diagonal-pivot panels as tall recursive
// factorisations of all their rows.
static bool block_panels() {
  const char* e = std::getenv("QUADESTH_BLOCK_PANEL");
  return !(e && e[0] == '0');
}

// Steps of a diagonal block with at least this many trailing tiles per side
// take the two-launch form (QUADESTH_BLK_SPLIT; 0: never).  The default 16
// keeps the 1024-column panels' steps (at most 15) fused, where the panel's
// chain of launches is the critical path, and splits the early steps of the
// wide panels, where the chip is full and the fused form's extra products
// cost the trailing update time.  Factor at 76,806 DOF, one box: never
// 421.8 ms, from 48 / 32 / 16 / 8 / 4 / 1 tiles 416.1 / 413.8 / 410.9 / 413.4 /
// 415.6 / 413.8 ms.
// QUADESTH_INC_INVERSE=0 forms U11^-1 and L11^-1 by recursive doubling in
// every panel.
static bool inc_inverses() {
  const char* e = std::getenv("QUADESTH_INC_INVERSE");
  return !(e && e[0] == '0');
}

static int blk_split_nt() {
  const char* e = std::getenv("QUADESTH_BLK_SPLIT");
  return e ? std::atoi(e) : 16;
}

// Diagonal-pivot panel [j, j+jb) with rows below it (jb a multiple of 64):
// the jb x jb diagonal block factored in 64-column steps of one launch each,
// U11^-1 (whose diagonal blocks the steps leave in Y) completed by recursive
// doubling, then L21 = A21 U11^-1 in one FP16 product into W.  The growth of
// L21 is taken by the trailing update's conversion of it (update_cols).
static int panel_block(LuCtx<float>& c, long long j, int jb) {
  qd_lu_state* s = c.s;
  const int npad = binv_pad(jb), T = jb / 64;
  const long long r0 = j + jb, M = c.N - r0;
  const int Kp = f16_kpad(jb);
  const long long nb = (jb + 63) / 64;
  qd_lu_state::F16Scratch& sc = s->sc_l21;
  if (jb % 64 != 0 || T < 2)
    return set_error(c.h, QD_ERR_STATE, "diagonal-block panel of %d columns (a multiple of 64 from 128 up)", jb);
  if (s->binv.count < (size_t)npad * npad || s->binvl.count < (size_t)npad * npad ||
      s->binv_t.count < (size_t)npad * npad / 2 ||
      s->bdiag.count < (size_t)2 * npad || sc.h16a.count < (size_t)(M * Kp) || sc.h16b.count < (size_t)jb * Kp ||
      sc.colscale.count < (size_t)nb || s->bstep.count < (size_t)(6 * T + 4) * 4096)
    return set_error(c.h, QD_ERR_STATE, "diagonal-block panel buffers sized for fewer than %d columns", jb);
  float* Y = s->binv.p;
  float* Tm = s->binv_t.p;
  float* dg = s->bdiag.p;
  float* linv = s->ts_linv.p;
  // the diagonal block: tile 0 (with the copies of tile column and row 0),
  // then one launch per step
  float* Z = s->binvl.p;
  BlkBufs bb{s->bstep.p, s->bstep.p + (size_t)2 * T * 4096, s->bstep.p + (size_t)4 * T * 4096, T, Z,
             s->bstep.p + (size_t)(4 * T + 4) * 4096, s->bstep.p + (size_t)(5 * T + 4) * 4096};
  const int split_nt = blk_split_nt();
  hipLaunchKernelGGL(k_blk_step, dim3((unsigned)(2 * T - 1)), dim3(256), 0, c.st, c.W, c.ld, j, -1, 0, bb, Y, npad, dg,
                     linv, s->ipiv_dev.p, s->growth.p);
  // a panel whose steps all take one launch forms U11^-1 and L11^-1 in them
  // (tile column and row k in launch k, the last ones in a launch of their
  // own) instead of by recursive doubling afterwards
  const bool inc = inc_inverses() && !(split_nt > 0 && T - 1 >= split_nt);
  for (int k = 0; k + 1 < T; ++k) {
    const int nt = T - 1 - k;
    if (split_nt > 0 && nt >= split_nt) {
      hipLaunchKernelGGL(k_blk_solve2, dim3((unsigned)(2 * nt)), dim3(256), 0, c.st, c.W, c.ld, j, k, nt, bb,
                         s->growth.p);
      hipLaunchKernelGGL(k_blk_update2, dim3((unsigned)(nt * nt)), dim3(256), 0, c.st, c.W, c.ld, j, k, nt, bb, Y,
                         npad, dg, linv, s->ipiv_dev.p, s->growth.p);
    } else {
      hipLaunchKernelGGL(k_blk_step, dim3((unsigned)(nt * nt + (inc ? 2 * k : 0))), dim3(256), 0, c.st, c.W, c.ld, j,
                         k, nt, bb, Y, npad, dg, linv, s->ipiv_dev.p, s->growth.p);
    }
  }
  if (inc)
    hipLaunchKernelGGL(k_blk_step, dim3((unsigned)(2 * (T - 1))), dim3(256), 0, c.st, c.W, c.ld, j, T - 1, 0, bb, Y,
                       npad, dg, linv, s->ipiv_dev.p, s->growth.p);
  QD_HIP(c.h, hipGetLastError());
  if (npad > jb && !inc)
    hipLaunchKernelGGL(k_binv_identity, dim3((unsigned)(npad / 64 - T)), dim3(256), 0, c.st, T, npad, Y, Z);
  // U11^-1 and L11^-1 by recursive doubling, both in each launch.  Pair p
  // covers rows and columns [2pb, 2pb + 2b): T_p = U12_p Y2_p, Y12_p =
  // -Y1_p T_p (Y upper triangular), and T'_p = L21_p Z1_p, Z21_p = -Z2_p T'_p
  // (Z lower triangular)
  const float* Wd = c.W + j * c.ld + j;
  constexpr int kNoLim = 1 << 30;
  float* Tl = Tm + (size_t)npad * npad / 4;
  for (int b = 64; b < npad && !inc; b *= 2) {
    const int P = npad / (2 * b);
    const long long sy = 2LL * b * (npad + 1), su = 2LL * b * (c.ld + 1), st = (long long)b * b;
    const long long bn = (long long)b * npad;
    const BatOp u1{Wd + b, c.ld, su, 0, b, 2 * b, jb, 0, Y + bn + b, (long long)npad, sy, 1, Tm, (long long)b, st, 1.0f};
    const BatOp l1{Wd + (long long)b * c.ld, c.ld, su, b, 0, 2 * b, jb, 0, Z, (long long)npad, sy, 2, Tl, (long long)b,
                   st, 1.0f};
    const BatOp u2{Y, (long long)npad, sy, 0, 0, 0, kNoLim, 1, Tm, (long long)b, st, 0, Y + b, (long long)npad, sy,
                   -1.0f};
    const BatOp l2{Z + bn + b, (long long)npad, sy, 0, 0, 0, kNoLim, 2, Tl, (long long)b, st, 0, Z + bn,
                   (long long)npad, sy, -1.0f};
    // 64 x 64 tiles at every level: 4x the workgroups of 128 x 128 ones,
    // which the few large products of the last levels need
    const dim3 g((unsigned)(b / 64), (unsigned)(b / 64), (unsigned)(2 * P));
    hipLaunchKernelGGL(k_gemm_bat<64>, g, dim3(256), 0, c.st, b, P, u1, l1);
    hipLaunchKernelGGL(k_gemm_bat<64>, g, dim3(256), 0, c.st, b, P, u2, l2);
  }
  {
    // L11^-1's strictly lower part, negated, in FP16 for the U12 solves
    const int lb = (s->linvh_col[0] == -1 || (s->linvh_col[1] != -1 && s->linvh_col[0] < s->linvh_col[1])) ? 0 : 1;
    if (s->linvh[lb].count >= (size_t)jb * Kp) {
      hipLaunchKernelGGL(k_linv_a16, dim3((unsigned)(((long long)jb * (Kp / 8) + 255) / 256)), dim3(256), 0, c.st, jb,
                         npad, Kp, (const float*)Z, s->linvh[lb].p);
      s->linvh_col[lb] = j;
      s->linvh_w[lb] = jb;
    }
  }
  hipLaunchKernelGGL(k_binv_rowscale, dim3((unsigned)(((long long)jb * (npad / 4) + 255) / 256)), dim3(256), 0, c.st,
                     jb, npad, Y, (const float*)dg);
  hipLaunchKernelGGL(k_scale_cvt_b16t, dim3((unsigned)nb), dim3(256), 0, c.st, jb, (long long)jb, Kp, (const float*)Y,
                     (long long)npad, sc.colscale.p, sc.h16b.p, 1);
  float* A21 = c.W + r0 * c.ld + j;
  const int va = (c.ld % 4 == 0 && j % 4 == 0 && jb % 8 == 0) ? 1 : 0;
  hipLaunchKernelGGL(k_cvt_a16_l21, dim3((unsigned)((M * (Kp / 8) + 255) / 256)), dim3(256), 0, c.st, M, jb, Kp, A21,
                     c.ld, (const float*)(dg + npad), sc.h16a.p, va);
  QD_HIP(c.h, hipGetLastError());
  const int rows16 = (c.ld % 4 == 0 && (reinterpret_cast<uintptr_t>(A21) & 15) == 0) ? 1 : 0;
  // the product also leaves L21 in FP16 for the trailing updates, with its
  // growth (the panels alternate between two copies: the main stream still
  // reads the previous panel's while this one is written)
  const int b = s->l21h_next;
  Out16 o16{nullptr, 0, nullptr};
  if (s->l21h[b].count >= (size_t)(M * Kp)) {
    o16 = Out16{s->l21h[b].p, (long long)Kp, s->growth.p};
    s->l21h_col[b] = j;
    s->l21h_next = b ^ 1;
  }
  // N is strictly upper triangular: a tile's K runs below its last column
  return launch_f16_update(c.h, false, M, jb, Kp, sc.h16a.p, sc.h16b.p, sc.colscale.p, A21, c.ld, rows16, 0, c.st,
                           o16, 1);
}

// The FP16 L21 of the panel at column j, if its product wrote one.
static const _Float16* l21_copy(const qd_lu_state* s, long long j) {
  for (int b = 0; b < 2; ++b)
    if (s->l21h_col[b] == j) return s->l21h[b].p;
  return nullptr;
}

// Panel of columns [j, j+jb) on the context's stream, timed as one lu_panel.
template <typename T>
static int panel_step(LuCtx<T>& c, long long j, int jb) {
  const double m = (double)(c.N - j);
  ProfScope ps(c.h, QD_PROF_LU_PANEL, m * jb * (double)jb - (double)jb * jb * jb / 3.0, c.st);
  if constexpr (std::is_same<T, float>::value) {
    if (c.calu && c.diag && c.f16 && block_panels() && jb > kDW && jb % kDW == 0 && j % 4 == 0 && j + jb < c.N &&
        c.ld % 4 == 0)
      return panel_block(c, j, jb);
  }
  return panel(c, j, jb, j, j + jb);
}

// A22[:, c0:c1) -= L21 U12[:, c0:c1) for the step whose panel is [j, j+jb).
template <typename T>
static int update_cols(LuCtx<T>& c, long long j, int jb, long long c0, long long c1, bool reuse_a16,
                       int max_wgs = 0) {
  const long long r0 = j + jb;
  // with diagonal pivots the first conversion of L21 also takes its growth
  unsigned* g = (c.calu && c.diag && !reuse_a16) ? c.s->growth.p : nullptr;
  return gemm_sub(c, c.N - r0, c1 - c0, jb, c.W + r0 * c.ld + j, c.W + j * c.ld + c0, c.W + r0 * c.ld + c0, true,
                  true, reuse_a16, max_wgs, g);
}

// Largest multiplier |l| with which a diagonal-pivot factorisation is kept.
// At 1 every diagonal pivot is the largest |a| of its column below it, i.e.
// the pivot partial pivoting (LAPACK's first largest) would have taken, so a
// kept factorisation has partial pivoting's pivot sequence under this
// factorisation's own rounding.
constexpr float kDiagMaxMult = 1.0f;

// QUADESTH_PIVOT=tournament skips the diagonal-pivot attempt.
static bool diag_pivots_allowed() {
  const char* e = std::getenv("QUADESTH_PIVOT");
  return !(e && std::strcmp(e, "tournament") == 0);
}

// The FP16 trailing update of columns [c0, c1) in two launches: the next
// panel's rows [j+jb, j+jb+split) first (ev recorded after it), then the rest.
// L21's FP16 copy is already in c.sc (converted by the update of the next
// panel's columns on the same stream).
template <typename T>
static int update_split(LuCtx<T>& c, long long j, int jb, long long c0, long long c1, long long split,
                        hipEvent_t ev) {
  const long long r0 = j + jb, M = c.N - r0, Nn = c1 - c0;
  const int K = jb, Kp = f16_kpad(K);
  const float* B = reinterpret_cast<const float*>(c.W + j * c.ld + c0);
  const long long nb = (Nn + 63) / 64;
  QD_ALLOC(c.h, c.sc->colscale, (size_t)nb);
  QD_ALLOC(c.h, c.sc->h16b, (size_t)Nn * Kp);
  const int vb = ((reinterpret_cast<uintptr_t>(B) & 15) == 0 && c.ld % 4 == 0) ? 1 : 0;
  hipLaunchKernelGGL(k_scale_cvt_b16t, dim3((unsigned)nb), dim3(256), 0, c.st, K, Nn, Kp, B, c.ld, c.sc->colscale.p,
                     c.sc->h16b.p, vb);
  float* C = reinterpret_cast<float*>(c.W + r0 * c.ld + c0);
  const int rows16 = (c.ld % 4 == 0 && (reinterpret_cast<uintptr_t>(C) & 15) == 0) ? 1 : 0;
  const _Float16* a16 = c.a16 ? c.a16 : c.sc->h16a.p;
  int rc;
  {
    ProfScope ps(c.h, QD_PROF_LU_TRAILING_GEMM, 2.0 * (double)split * (double)Nn * K, c.st);
    if ((rc = launch_f16_update(c.h, true, split, Nn, Kp, a16, c.sc->h16b.p, c.sc->colscale.p, C, c.ld, rows16, 0,
                                c.st)))
      return rc;
  }
  QD_HIP(c.h, hipEventRecord(ev, c.st));
  {
    ProfScope ps(c.h, QD_PROF_LU_TRAILING_GEMM, 2.0 * (double)(M - split) * (double)Nn * K, c.st);
    if ((rc = launch_f16_update(c.h, true, M - split, Nn, Kp, a16 + (size_t)split * Kp, c.sc->h16b.p,
                                c.sc->colscale.p, C + split * c.ld, c.ld, rows16, 0, c.st)))
      return rc;
  }
  QD_HIP(c.h, hipGetLastError());
  return QD_OK;
}

// Workgroup cap of the trailing update that runs beside a panel
// (QUADESTH_GEMM_WGS; 0 = one workgroup per tile).  Measured on MI355X at
// 76,806 DOF, caps of 320-512 did not shorten the factorisation: the panel
// beside the update is slowed by the memory system, not by CU slots.  Nor did
// CU-masked streams that kept the update off 16-64 CUs: the panel gained
// 9-21%, the update lost about as much time with the CUs.
static int lookahead_gemm_wgs() {
  if (const char* e = std::getenv("QUADESTH_GEMM_WGS")) return std::max(0, std::atoi(e));
  return 0;
}

// Panel widths of a factorisation: `wide` columns while more than `wide_rem`
// columns remain (wide = 0: never), `nb` after.  A wide panel halves the
// passes of the trailing update over C where that update hides the panel; the
// narrow one keeps the late, panel-bound steps short.
struct PanelWidths {
  int nb = 1024;
  int wide = 0;
  long long wide_rem = 0;
  int tail = 0;  // columns of the panels once at most tail_rem remain (0: nb)
  long long tail_rem = 0;
  int at(long long j, long long N) const {
    const long long rem = N - j;
    const int w = (wide > nb && rem > wide_rem) ? wide : (tail > 0 && rem <= tail_rem) ? tail : nb;
    return (int)std::min<long long>(w, rem);
  }
  int widest() const { return std::max(std::max(nb, wide), tail); }
};

template <typename T>
static int factor_steps(LuCtx<T>& c, const PanelWidths& pw, bool look);

// Blocked right-looking LU with a look-ahead of one panel: each step solves
// for and updates the next panel's columns first; that panel is then factored
// on the aux stream while the rest of U12 and of the trailing matrix is
// updated on the main stream (the two touch disjoint columns and separate
// FP16 scratch).  The main stream waits for the panel before the next step's
// interchanges.
// QUADESTH_LOOKAHEAD=0 runs the steps strictly in sequence.
template <typename T>
static int factor(LuCtx<T>& c, const PanelWidths& pw) {
  qd_lu_state* s = c.s;
  const int NB = pw.widest();
  const long long N = c.N;
  const size_t maxblocks = (size_t)((N + kColRowsPerBlock - 1) / kColRowsPerBlock) + 1;
  QD_ALLOC(c.h, s->piv_part, maxblocks);
  QD_ALLOC(c.h, s->piv_idx, maxblocks);
  QD_ALLOC(c.h, s->piv_state, 2);
  QD_ALLOC(c.h, s->rpiv, 1);
  QD_ALLOC(c.h, s->ipiv_dev, (size_t)N);
  if (c.calu) {
    long long ncand = 0, ncnt = 0;
    ts_tree_sizes(N, ncand, ncnt);
    const bool fresh = (long long)s->ts_cnt.count < ncnt + 1 || !s->ts_cnt.p;
    QD_ALLOC(c.h, s->ts_cand, (size_t)ncand);
    QD_ALLOC(c.h, s->ts_cnt, (size_t)ncnt + 1);
    QD_ALLOC(c.h, s->ts_meta, (size_t)kTsMeta);
    QD_ALLOC(c.h, s->ts_lu, (size_t)kDW * kDW);
    QD_ALLOC(c.h, s->ts_top, (size_t)kTsW * kTsW);
    QD_ALLOC(c.h, s->ts_linv, (size_t)((N + kTsW - 1) / kTsW + 2) * kTsW * kTsW);
    // the strip solves keep their strips in dynamic LDS
    QD_HIP(c.h, hipFuncSetAttribute(reinterpret_cast<const void*>(&k_trsm_strip<32>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 256 * 32 * (int)sizeof(float)));
    QD_HIP(c.h, hipFuncSetAttribute(reinterpret_cast<const void*>(&k_trsm_strip<8>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    kStripMaxRows * 8 * (int)sizeof(float)));
    if (fresh) QD_HIP(c.h, hipMemsetAsync(s->ts_cnt.p, 0, s->ts_cnt.count * sizeof(unsigned), c.st));
    QD_ALLOC(c.h, s->growth, 1);
    QD_HIP(c.h, hipMemsetAsync(s->growth.p, 0, sizeof(unsigned), c.st));
  }
  const char* la = std::getenv("QUADESTH_LOOKAHEAD");
  const bool look = !(la && la[0] == '0');
  if (c.f16) {
    // sized up front, so that nothing is reallocated while two streams run
    const size_t Kp = (size_t)f16_kpad(NB), Kh = (size_t)f16_kpad(NB / 2);
    QD_ALLOC(c.h, s->sc_main.h16a, (size_t)N * Kp);
    QD_ALLOC(c.h, s->sc_main.h16b, (size_t)N * Kp);
    QD_ALLOC(c.h, s->sc_main.colscale, (size_t)(N / 64 + 2));
    QD_ALLOC(c.h, s->sc_panel.h16a, (size_t)N * Kh);
    QD_ALLOC(c.h, s->sc_panel.h16b, Kh * Kh);
    QD_ALLOC(c.h, s->sc_panel.colscale, (size_t)(NB / 64 + 2));
    QD_ALLOC(c.h, s->sc_solve.h16a, Kh * Kh);
    // (B operands of the U12 solves through L11^-1 take K = the panel width)
    QD_ALLOC(c.h, s->sc_solve.h16b, (size_t)N * Kp);
    QD_ALLOC(c.h, s->sc_solve.colscale, (size_t)(N / 64 + 2));
    if (c.calu && c.diag && block_panels()) {
      const size_t np = (size_t)binv_pad(NB);
      QD_ALLOC(c.h, s->binv, np * np);
      QD_ALLOC(c.h, s->binvl, np * np);
      QD_ALLOC(c.h, s->binv_t, np * np / 2);
      QD_ALLOC(c.h, s->linvh[0], (size_t)NB * Kp);
      QD_ALLOC(c.h, s->linvh[1], (size_t)NB * Kp);
      QD_ALLOC(c.h, s->bdiag, 2 * np);
      QD_ALLOC(c.h, s->bstep, (size_t)(6 * (NB / 64) + 4) * 4096);
      QD_ALLOC(c.h, s->l21h[0], (size_t)N * Kp);
      QD_ALLOC(c.h, s->l21h[1], (size_t)N * Kp);
      QD_ALLOC(c.h, s->sc_l21.h16a, (size_t)N * Kp);
      QD_ALLOC(c.h, s->sc_l21.h16b, (size_t)NB * Kp);
      QD_ALLOC(c.h, s->sc_l21.colscale, (size_t)(NB / 64 + 2));
    }
    if (look) {
      QD_ALLOC(c.h, s->sc_u12.h16a, Kh * Kh);
      QD_ALLOC(c.h, s->sc_u12.h16b, (size_t)N * Kp);
      QD_ALLOC(c.h, s->sc_u12.colscale, (size_t)(N / 64 + 2));
    }
  }
  if (look && !s->aux) {
    QD_HIP(c.h, hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking));
    QD_HIP(c.h, hipEventCreateWithFlags(&s->ev_cols, hipEventDisableTiming));
    QD_HIP(c.h, hipEventCreateWithFlags(&s->ev_panel, hipEventDisableTiming));
    QD_HIP(c.h, hipEventCreateWithFlags(&s->ev_top, hipEventDisableTiming));
    QD_HIP(c.h, hipEventCreateWithFlags(&s->ev_u12, hipEventDisableTiming));
  }
  return factor_steps(c, pw, look);
}

template <typename T>
static int factor_steps(LuCtx<T>& c, const PanelWidths& pw, bool look) {
  qd_lu_state* s = c.s;
  const long long N = c.N;
  c.sc = &s->sc_main;
  c.sc_solve = &s->sc_solve;
  LuCtx<T> cp = c;
  cp.sc_solve = nullptr;
  cp.st = s->aux;
  cp.sc = &s->sc_panel;
  // With diagonal pivots (no interchanges) and FP16 updates, a step's big
  // update goes in two launches, the next panel's rows first; the aux stream
  // then solves the next step's U12 (beyond the panel after next) beside the
  // rest of the update, once the next panel is factored.
  const bool split = look && c.diag && c.f16;
  // The next U12 solve on the aux stream.  It can still run when the main
  // stream starts the next step's solve for the panel after next (whose
  // inner FP16 updates, with panels wider than the strip solve takes, use
  // sc_solve), so it has scratch of its own.
  LuCtx<T> ca = cp;
  ca.sc_solve = &s->sc_u12;
  bool u12_ready = false;  // this step's U12 beyond the next panel is solved (ev_u12)
  QD_HIP(c.h, hipMemsetAsync(s->piv_state.p, 0, 2 * sizeof(int), c.st));
  s->l21h_col[0] = s->l21h_col[1] = -1;  // no FP16 L21 or L11^-1 of this factorisation yet
  s->linvh_col[0] = s->linvh_col[1] = -1;
  int rc;
  const int jb0 = pw.at(0, N);
  if (c.src) {
    // W from A: the first panel's columns, then the rest beside that panel
    // (on the aux stream with the look-ahead; the two touch disjoint columns)
    // (also the row sums of |A| per 512-column slice: the first launch
    // covers whole slices, at least the first panel's columns)
    const unsigned nt = (unsigned)((N + 31) / 32);
    const unsigned ct = std::min(nt, (unsigned)((jb0 + 32 * kAbsTiles - 1) / (32 * kAbsTiles)) * kAbsTiles);
    const unsigned nsl = (nt + kAbsTiles - 1) / kAbsTiles;
    QD_ALLOC(c.h, s->rowabs, (size_t)nsl * N + N);
    auto wg = [](unsigned tiles) { return (tiles + kAbsTiles - 1) / kAbsTiles; };
    hipLaunchKernelGGL(k_transpose_abs<T>, dim3(wg(ct), nt), dim3(256), 0, c.st, N, c.src, c.src_ld, c.W, c.ld, 0LL,
                       s->rowabs.p);
    QD_HIP(c.h, hipGetLastError());
    if (look) QD_HIP(c.h, hipEventRecord(s->ev_cols, c.st));
    // the rest is enqueued before the panel's many small launches, which
    // would hold the host up while the main stream idles
    if (nt > ct)
      hipLaunchKernelGGL(k_transpose_abs<T>, dim3(wg(nt - ct), nt), dim3(256), 0, c.st, N, c.src, c.src_ld, c.W, c.ld,
                         (long long)ct, s->rowabs.p);
    QD_HIP(c.h, hipGetLastError());
    if (look) {
      QD_HIP(c.h, hipStreamWaitEvent(cp.st, s->ev_cols, 0));
      if ((rc = panel_step(cp, 0, jb0))) return rc;
      QD_HIP(c.h, hipEventRecord(s->ev_panel, cp.st));
      QD_HIP(c.h, hipStreamWaitEvent(c.st, s->ev_panel, 0));
    } else if ((rc = panel_step(c, 0, jb0))) {
      return rc;
    }
  } else if ((rc = panel_step(c, 0, jb0))) {
    return rc;
  }
  for (long long j = 0, step = 0; j < N; j += step) {
    const int jb = pw.at(j, N);
    step = jb;
    c.a16 = c.f16 ? l21_copy(s, j) : nullptr;
    const long long nj = j + jb;
    const int njb = nj < N ? pw.at(nj, N) : 0;
    // interchanges on the columns left and right of the panel (none with
    // diagonal pivots)
    if (j > 0 && !c.diag)
      hipLaunchKernelGGL(k_laswp<T>, dim3((unsigned)((j + 255) / 256)), dim3(256), 0, c.st, c.W, c.ld, 0LL,
                         j, (int)j, (int)nj, s->ipiv_dev.p);
    if (nj < N) {
      if (!c.diag)
        hipLaunchKernelGGL(k_laswp<T>, dim3((unsigned)((N - nj + 255) / 256)), dim3(256), 0, c.st, c.W, c.ld, nj, N,
                           (int)j, (int)nj, s->ipiv_dev.p);
      QD_HIP(c.h, hipGetLastError());
      if (look && nj + njb < N) {
        // the next panel's columns first: U12 and its update, then the panel
        // can start on the aux stream
        if ((rc = trsm_lower(c, j, jb, nj, nj + njb))) return rc;
        if ((rc = update_cols(c, j, jb, nj, nj + njb, false))) return rc;
        QD_HIP(c.h, hipEventRecord(s->ev_cols, c.st));
        // The rest of the step is enqueued before the panel: the host can
        // be held up enqueuing the panel's many small launches, and the
        // main stream would idle meanwhile.
        if (u12_ready) {
          QD_HIP(c.h, hipStreamWaitEvent(c.st, s->ev_u12, 0));
        } else if ((rc = trsm_lower(c, j, jb, nj + njb, N))) {
          return rc;
        }
        // L21's FP16 copy from the update above is still in sc_main (the U12
        // solve's inner updates use sc_solve)
        if (split) {
          if ((rc = update_split(c, j, jb, nj + njb, N, njb, s->ev_top))) return rc;
        } else if ((rc = update_cols(c, j, jb, nj + njb, N, true, lookahead_gemm_wgs()))) {
          return rc;
        }
        QD_HIP(c.h, hipStreamWaitEvent(cp.st, s->ev_cols, 0));
        if ((rc = panel_step(cp, nj, njb))) return rc;
        QD_HIP(c.h, hipEventRecord(s->ev_panel, cp.st));
        u12_ready = false;
        if (split) {
          // the next step's U12 beyond its next panel, when that step takes
          // this branch too: its rows are final once the first launch above
          // is done, and its L11 once the panel is
          const long long nj2 = nj + njb;
          const int njb2 = nj2 < N ? pw.at(nj2, N) : 0;
          if (nj2 < N && nj2 + njb2 < N) {
            QD_HIP(c.h, hipStreamWaitEvent(ca.st, s->ev_top, 0));
            if ((rc = trsm_lower(ca, nj, njb, nj2 + njb2, N))) return rc;
            QD_HIP(c.h, hipEventRecord(s->ev_u12, ca.st));
            u12_ready = true;
          }
        }
        QD_HIP(c.h, hipStreamWaitEvent(c.st, s->ev_panel, 0));
      } else {
        if (u12_ready) return set_error(c.h, QD_ERR_STATE, "factor: U12 solved ahead for a step without look-ahead");
        if ((rc = trsm_lower(c, j, jb, nj, N))) return rc;
        if ((rc = update_cols(c, j, jb, nj, N, false))) return rc;
        if (njb > 0 && (rc = panel_step(c, nj, njb))) return rc;
      }
    }
  }
  return QD_OK;
}

// ---------------------------------------------------------------------------
// triangular solves with row-major factors, nrhs vectors of stride N
constexpr int kTB = 64;

// Solve the kTB-block [i0, i0+m) of L (unit, lower=1) or U (upper, lower=0)
// in place on x; one workgroup.
template <typename T>
__global__ __launch_bounds__(256) void k_trsv_block(const T* __restrict__ W, long long ld, long long N, long long i0,
                                                    int m, int lower, int nrhs, T* __restrict__ x) {
  __shared__ T sL[kTB][kTB + 1];
  __shared__ T sx[8][kTB];
  for (int i = threadIdx.x; i < m * m; i += 256) {
    const int a = i / m, b = i % m;
    sL[a][b] = W[(i0 + a) * ld + i0 + b];
  }
  for (int i = threadIdx.x; i < m * nrhs; i += 256) sx[i / m][i % m] = x[(long long)(i / m) * N + i0 + i % m];
  __syncthreads();
  const int r = threadIdx.x % kTB, q = threadIdx.x / kTB;  // q: rhs lane group (4)
  if (lower) {
    for (int k = 0; k < m; ++k) {
      if (r > k && r < m)
        for (int v = q; v < nrhs; v += 4) sx[v][r] -= sL[r][k] * sx[v][k];
      __syncthreads();
    }
  } else {
    for (int k = m - 1; k >= 0; --k) {
      if (r == k)
        for (int v = q; v < nrhs; v += 4) sx[v][k] = sx[v][k] / sL[k][k];
      __syncthreads();
      if (r < k)
        for (int v = q; v < nrhs; v += 4) sx[v][r] -= sL[r][k] * sx[v][k];
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < m * nrhs; i += 256) x[(long long)(i / m) * N + i0 + i % m] = sx[i / m][i % m];
}

// x[rows] -= W[rows, cols] x[cols] for rows [r0, r1), cols [c0, c0+m), m <= kTB.
template <typename T>
__global__ __launch_bounds__(256) void k_trsv_update(const T* __restrict__ W, long long ld, long long N, long long r0,
                                                     long long r1, long long c0, int m, int nrhs, T* __restrict__ x) {
  __shared__ T sW[64][kTB + 1];
  __shared__ T sx[8][kTB];
  const long long rb = r0 + (long long)blockIdx.x * 64;
  for (int i = threadIdx.x; i < m * nrhs; i += 256) sx[i / m][i % m] = x[(long long)(i / m) * N + c0 + i % m];
  for (int i = threadIdx.x; i < 64 * m; i += 256) {
    const int a = i / m, b = i % m;
    sW[a][b] = (rb + a < r1) ? W[(rb + a) * ld + c0 + b] : (T)0;
  }
  __syncthreads();
  const int a = threadIdx.x % 64, q = threadIdx.x / 64;
  if (rb + a >= r1) return;
  for (int v = q; v < nrhs; v += 4) {
    T s = 0;
    for (int b = 0; b < m; ++b) s += sW[a][b] * sx[v][b];
    x[(long long)v * N + rb + a] -= s;
  }
}

// ---- triangular solves with inverted diagonal blocks ------------------------
// After a factorisation the TB x TB diagonal blocks of L (unit lower) and U
// are inverted once; every solve then advances one block per step with a
// parallel block product (x_I = inv(D_I) x_I) and an update of the remaining
// rows, instead of TB sequential substitution steps.
template <typename T>
struct TriTB {
  static constexpr int v = sizeof(T) == 4 ? 128 : 64;
};

// One workgroup of TB threads per diagonal block; thread j forms column j of
// the inverse by substitution on e_j.  inv: [block][TB][TB] row-major, zero
// padded beyond the matrix.
template <typename T, int TB>
__global__ __launch_bounds__(TB) void k_diag_inverse(const T* __restrict__ W, long long ld, long long N, int lower,
                                                     T* __restrict__ inv) {
  __shared__ T sA[TB][TB + 1];
  __shared__ T sX[TB][TB + 1];
  const long long i0 = (long long)blockIdx.x * TB;
  const int m = (int)std::min<long long>(TB, N - i0);
  const int j = threadIdx.x;
  for (int i = 0; i < TB; ++i) sA[i][j] = (i < m && j < m) ? W[(i0 + i) * ld + i0 + j] : (T)0;
  __syncthreads();
  if (lower) {
    for (int i = 0; i < TB; ++i) {
      T s = (i == j) ? (T)1 : (T)0;
      for (int k = 0; k < i; ++k) s -= sA[i][k] * sX[k][j];
      sX[i][j] = (i < m && j < m) ? s : (T)0;
    }
  } else {
    for (int i = TB - 1; i >= 0; --i) {
      T s = (i == j) ? (T)1 : (T)0;
      for (int k = i + 1; k < m; ++k) s -= sA[i][k] * sX[k][j];
      sX[i][j] = (i < m && j < m) ? s / sA[i][i] : (T)0;
    }
  }
  T* out = inv + (size_t)blockIdx.x * TB * TB;
  for (int i = 0; i < TB; ++i) out[i * TB + j] = sX[i][j];
}

// x[v][rows] -= W[rows, c0:c0+m] x[v][c0:c0+m] for rows [r0, r1), m <= TB,
// TB * sizeof(T) = 512 bytes.  32 lanes per row, each with one 16-byte
// piece of it and of x; the row sum is reduced across the 32 lanes.  A
// workgroup of 256 threads covers 8 rows per pass and kBuRows rows.  Columns
// at or beyond m (the partial last block) are masked before loading, since
// they may lie past the end of the row.
constexpr int kBuRows = 64, kBuMaxRhs = 8;

// x[v][rbase .. rbase+kBuRows) (rows below rend skipped) -= W[rows, c0:c0+m]
// x[v][c0:c0+m], with xr the lane's piece of x[v][c0:c0+m].
template <typename T, int TB>
__device__ __forceinline__ void bu_rows(const T* __restrict__ W, long long ld, long long N, long long rbase,
                                        long long rend, long long c0, int m, int nrhs, T* __restrict__ x,
                                        const T (&xr)[kBuMaxRhs][16 / sizeof(T)]) {
  constexpr int E = 16 / sizeof(T);  // elements per lane
  const int sub = threadIdx.x & 31, slot = threadIdx.x >> 5;
  const int col = sub * E;
  const bool full = (m == TB);
  constexpr int kPass = kBuRows / 8;
  // all rows' loads first, then the reductions
  T w[kPass][E];
#pragma unroll
  for (int pass = 0; pass < kPass; ++pass) {
    const long long r = rbase + pass * 8 + slot;
    const T* wr = W + r * ld + c0 + col;
#pragma unroll
    for (int e = 0; e < E; ++e) w[pass][e] = (T)0;
    if (r < rend) {
      if (full) {
        if constexpr (sizeof(T) == 4) {
          const float4 q = *reinterpret_cast<const float4*>(wr);
          w[pass][0] = q.x;
          w[pass][1] = q.y;
          w[pass][2] = q.z;
          w[pass][3] = q.w;
        } else {
          const double2 q = *reinterpret_cast<const double2*>(wr);
          w[pass][0] = q.x;
          w[pass][1] = q.y;
        }
      } else {
#pragma unroll
        for (int e = 0; e < E; ++e)
          if (col + e < m) w[pass][e] = wr[e];
      }
    }
  }
#pragma unroll
  for (int v = 0; v < kBuMaxRhs; ++v) {
    if (v < nrhs) {
#pragma unroll
      for (int pass = 0; pass < kPass; ++pass) {
        T sum = 0;
#pragma unroll
        for (int e = 0; e < E; ++e) sum += w[pass][e] * xr[v][e];
#pragma unroll
        for (int off = 16; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 32);
        const long long r = rbase + pass * 8 + slot;
        if (sub == 0 && r < rend) x[(long long)v * N + r] -= sum;
      }
    }
  }
}

// One block step of a triangular solve, one launch: every workgroup forms the
// block's solution x_b = inv_b y_b itself (y_b = the block's rows of the
// working vector, final once all earlier blocks' updates are in), then updates
// its share of rows [r0, r1) with it; workgroup 0 also writes x_b to the output
// vector.  The working vector's block rows are never written here, so no
// workgroup can read a block value another has already replaced.
template <typename T, int TB>
__global__ __launch_bounds__(256) void k_block_step(const T* __restrict__ W, long long ld, long long N, long long r0,
                                                    long long r1, long long c0, int m, const T* __restrict__ invb,
                                                    int nrhs, T* __restrict__ work, T* __restrict__ out) {
  static_assert((16 / sizeof(T)) * 32 == TB, "row piece layout");
  constexpr int E = 16 / sizeof(T);
  __shared__ T sy[kBuMaxRhs][TB];
  __shared__ T sxb[kBuMaxRhs][TB];
  for (int i = threadIdx.x; i < TB * nrhs; i += 256) {
    const int v = i / TB, r = i % TB;
    sy[v][r] = (r < m) ? work[(long long)v * N + c0 + r] : (T)0;
  }
  __syncthreads();
  // x_b: rows of the inverse read as 16-byte pieces, one (vector, row) per thread
  for (int i = threadIdx.x; i < m * nrhs; i += 256) {
    const int v = i / m, r = i % m;
    const T* row = invb + (size_t)r * TB;
    T acc = 0;
#pragma unroll
    for (int k0 = 0; k0 < TB; k0 += E) {
      T q[E];
      if constexpr (sizeof(T) == 4) {
        const float4 t = *reinterpret_cast<const float4*>(row + k0);
        q[0] = t.x;
        q[1] = t.y;
        q[2] = t.z;
        q[3] = t.w;
      } else {
        const double2 t = *reinterpret_cast<const double2*>(row + k0);
        q[0] = t.x;
        q[1] = t.y;
      }
#pragma unroll
      for (int e = 0; e < E; ++e) acc += q[e] * sy[v][k0 + e];
    }
    sxb[v][r] = acc;
    if (blockIdx.x == 0) out[(long long)v * N + c0 + r] = acc;
  }
  __syncthreads();
  T xr[kBuMaxRhs][E];
  const int col = (threadIdx.x & 31) * E;
#pragma unroll
  for (int v = 0; v < kBuMaxRhs; ++v)
#pragma unroll
    for (int e = 0; e < E; ++e) xr[v][e] = (v < nrhs && col + e < m) ? sxb[v][col + e] : (T)0;
  if (r1 > r0) bu_rows<T, TB>(W, ld, N, r0 + (long long)blockIdx.x * kBuRows, r1, c0, m, nrhs, work, xr);
}

// A whole triangular solve with one right-hand side in one launch ("decoupled
// look-back").  Block row b (b = i forward, nblk-1-i backward for the i-th
// group of kLbParts workgroups) is shared by kLbParts workgroups, which split
// its dependencies (the d-th earlier block in solve order, d = 0 .. ndep-1)
// round robin from the newest down: part p takes the d with
// (ndep-1-d) % kLbParts == kLbParts-1-p, so the last part holds d = ndep-1,
// the block solved just before.  Each part accumulates L[b, j] x_j as each x_j
// is published (a per-block flag set to this call's epoch).  Parts
// 0 .. kLbParts-2 publish their partial sums (flag per part); the last part
// collects them before its final dependency (they depend only on older
// blocks), then forms x_b = inv_b (y_b - sum) and publishes it: one
// publication per block on the critical path.  The long block rows near the
// end are streamed by several workgroups at once.
// A publication of partial sums is: the values' stores, a wait for their
// completion (vmcnt(0): __syncthreads alone waits only for LDS), a barrier,
// the flag.  For FP32, x_b is published in pub as 64-bit words (epoch in the
// high half, the value's bits in the low half), which readers poll directly:
// no flag, and no wait for the stores.  FP64 values take the flag route.
// A workgroup waits only on workgroups with smaller indices (earlier blocks,
// or earlier parts of its own block), which are dispatched before it.  The
// waits are bounded: after kLbMaxSpins polls (about 0.1 s) one gives up and
// sets *err to the epoch, and once *err holds the epoch no workgroup of the
// call waits any more (a wrong result, never a hang; the solve's caller reads
// *err and reports the failure).  Values and flags are relaxed agent-scope
// atomics (they bypass the caches that are not coherent across XCDs).  The last
// part loads the block's inverse before its first wait, and every part loads
// the W block of its next dependency before polling that dependency's flag.
constexpr int kLbMaxSpins = 1 << 17;
constexpr int kLbParts = 4;

// One bounded wait of thread 0 for flags[i] == epoch.
__device__ __forceinline__ void lb_wait(const int* flags, long long i, int epoch, int* err) {
  int spins = 0;
  while (__hip_atomic_load(&flags[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
    if (++spins > kLbMaxSpins) {
      atomicExch(err, epoch);
      break;
    }
    if ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch) break;
    __builtin_amdgcn_s_sleep(2);
  }
}

// flags: [nblk] block flags, then [nblk * kLbParts] part flags; part:
// partial sums, nblk * kLbParts * TB.
template <typename T, int TB>
__global__ __launch_bounds__(256) void k_trsv_lookback(const T* __restrict__ W, long long ld, long long N, int lower,
                                                       const T* __restrict__ inv, const T* __restrict__ y,
                                                       T* out, T* part, int* flags, unsigned long long* pub,
                                                       int epoch, int* err) {
  constexpr int E = 16 / sizeof(T);  // elements of a lane's row piece
  constexpr int kP = TB / 8;          // rows per thread (8 rows per pass)
  constexpr bool kPacked = sizeof(T) == 4;
  static_assert(E * 32 == TB, "row piece layout");
  __shared__ T sres[TB];
  __shared__ T sxj[2][TB];  // packed route: dependency values, double-buffered
  __shared__ T sy[TB];      // last part: the block's right-hand side, loaded at the start
  const long long nblk = (N + TB - 1) / TB;
  const long long grp = (long long)blockIdx.x / kLbParts;
  const int pidx = (int)(blockIdx.x % kLbParts);
  const bool last = (pidx == kLbParts - 1);
  const long long b = lower ? grp : nblk - 1 - grp;
  const long long rb = b * TB;
  const int m = (int)min((long long)TB, N - rb);
  const int sub = threadIdx.x & 31, slot = threadIdx.x >> 5;
  const int col = sub * E;
  int* pflags = flags + nblk;
  T acc[kP];
#pragma unroll
  for (int p = 0; p < kP; ++p) acc[p] = (T)0;
  // the block's inverse (last part only), loaded before any wait: thread t
  // holds part t % kParts of row t / kParts (kParts = 256 / TB parts of kIp)
  constexpr int kParts = 256 / TB, kIp = TB / kParts;
  static_assert(kIp % E == 0, "inverse parts are whole 16-byte pieces");
  const int irow = threadIdx.x / kParts, ipart = threadIdx.x % kParts;
  T pinv[kIp];
  if (last) {
    const T* src = inv + (size_t)b * TB * TB + (size_t)irow * TB + ipart * kIp;
#pragma unroll
    for (int k0 = 0; k0 < kIp; k0 += E) {
      if constexpr (sizeof(T) == 4) {
        const float4 q = *reinterpret_cast<const float4*>(src + k0);
        pinv[k0] = q.x;
        pinv[k0 + 1] = q.y;
        pinv[k0 + 2] = q.z;
        pinv[k0 + 3] = q.w;
      } else {
        const double2 q = *reinterpret_cast<const double2*>(src + k0);
        pinv[k0] = q.x;
        pinv[k0 + 1] = q.y;
      }
    }
    // y_b now, off the critical path (read after a later barrier)
    if (threadIdx.x < TB) sy[threadIdx.x] = ((int)threadIdx.x < m) ? y[rb + threadIdx.x] : (T)0;
  }
  const long long ndep = lower ? b : nblk - 1 - b;
  // the other parts' sums (last part, lanes with sub == 0: rows slot + 8 p)
  T psum[kP];
#pragma unroll
  for (int p = 0; p < kP; ++p) psum[p] = (T)0;
  auto gather = [&]() {
    if (threadIdx.x == 0)
      for (int q = 0; q < kLbParts - 1; ++q) lb_wait(pflags, b * kLbParts + q, epoch, err);
    __syncthreads();
    if (sub == 0) {
#pragma unroll
      for (int p = 0; p < kP; ++p) {
        const int r = slot + 8 * p;
        T sum = (T)0;
        for (int q = 0; q < kLbParts - 1; ++q)
          sum += __hip_atomic_load(&part[((size_t)b * kLbParts + q) * TB + r], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        psum[p] = sum;
      }
    }
  };
  bool gathered = false;
  const long long d0 = ((ndep - kLbParts + pidx) % kLbParts + kLbParts) % kLbParts;
  int buf = 0;
  for (long long d = d0; d < ndep; d += kLbParts) {
    const long long j = lower ? d : nblk - 1 - d;
    const long long cj = j * TB;
    const int mj = (int)min((long long)TB, N - cj);
    T w[kP][E];
#pragma unroll
    for (int p = 0; p < kP; ++p) {
      const int r = slot + 8 * p;
      const T* src = W + (rb + r) * ld + cj + col;
#pragma unroll
      for (int e = 0; e < E; ++e) w[p][e] = (T)0;
      if (r < m) {
        if (col + E <= mj) {
          if constexpr (sizeof(T) == 4) {
            // each L/U entry is read once per sweep: non-temporal
            typedef float f4v __attribute__((ext_vector_type(4)));
            const f4v q = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(src));
            w[p][0] = q.x;
            w[p][1] = q.y;
            w[p][2] = q.z;
            w[p][3] = q.w;
          } else {
            const double2 q = *reinterpret_cast<const double2*>(src);
            w[p][0] = q.x;
            w[p][1] = q.y;
          }
        } else {
#pragma unroll
          for (int e = 0; e < E; ++e)
            if (col + e < mj) w[p][e] = src[e];
        }
      }
    }
    if (last && d + kLbParts >= ndep) {
      gather();
      gathered = true;
    }
    T xj[E];
    if constexpr (kPacked) {
      // the first 32 lanes poll x_j's words (bounded), then share them in LDS;
      // the buffer alternates, so one barrier per dependency suffices
      if (slot == 0) {
        unsigned long long v[E];
        int spins = 0;
        for (;;) {
          bool ok = true;
#pragma unroll
          for (int e = 0; e < E; ++e) {
            v[e] = (col + e < mj) ? __hip_atomic_load(&pub[cj + col + e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                  : ((unsigned long long)(unsigned)epoch << 32);
            ok = ok && (unsigned)(v[e] >> 32) == (unsigned)epoch;
          }
          if (ok) break;
          if (++spins > kLbMaxSpins) {
            atomicExch(err, epoch);
            break;
          }
          if ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch) break;
          __builtin_amdgcn_s_sleep(1);
        }
#pragma unroll
        for (int e = 0; e < E; ++e) sxj[buf][col + e] = (col + e < mj) ? __uint_as_float((unsigned)v[e]) : 0.0f;
      }
      __syncthreads();
#pragma unroll
      for (int e = 0; e < E; ++e) xj[e] = sxj[buf][col + e];
      buf ^= 1;
    } else {
      if (threadIdx.x == 0) lb_wait(flags, j, epoch, err);
      __syncthreads();
#pragma unroll
      for (int e = 0; e < E; ++e)
        xj[e] = (col + e < mj) ? __hip_atomic_load(&out[cj + col + e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                               : (T)0;
    }
#pragma unroll
    for (int p = 0; p < kP; ++p)
#pragma unroll
      for (int e = 0; e < E; ++e) acc[p] += w[p][e] * xj[e];
  }
  // row sums across the 32 lanes
  T rs[kP];
#pragma unroll
  for (int p = 0; p < kP; ++p) {
    T sum = acc[p];
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 32);
    rs[p] = sum;
  }
  T* mypart = part + ((size_t)b * kLbParts + pidx) * TB;
  if (!last) {
    // publish this part's sums
#pragma unroll
    for (int p = 0; p < kP; ++p) {
      const int r = slot + 8 * p;
      if (sub == 0) __hip_atomic_store(&mypart[r], rs[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // the stores have completed before the flag is written: __syncthreads
    // alone does not wait for global stores
    __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0)
    __syncthreads();
    if (threadIdx.x == 0)
      __hip_atomic_store(&pflags[b * kLbParts + pidx], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  // the last part: the other parts' sums (collected before the final
  // dependency, or here when there was none), then the block's right-hand side
  if (!gathered) gather();
#pragma unroll
  for (int p = 0; p < kP; ++p) {
    const int r = slot + 8 * p;
    if (sub == 0) sres[r] = (r < m) ? sy[r] - (rs[p] + psum[p]) : (T)0;
  }
  __syncthreads();
  {
    T xb = 0;
#pragma unroll
    for (int k = 0; k < kIp; ++k) xb += pinv[k] * sres[ipart * kIp + k];
#pragma unroll
    for (int off = 1; off < kParts; off <<= 1) xb += __shfl_xor(xb, off);
    if constexpr (kPacked) {
      if (ipart == 0 && irow < m) {
        const unsigned long long w = ((unsigned long long)(unsigned)epoch << 32) | __float_as_uint(xb);
        __hip_atomic_store(&pub[rb + irow], w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        out[rb + irow] = xb;
      }
      return;
    } else {
      if (ipart == 0 && irow < m)
        __hip_atomic_store(&out[rb + irow], xb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): x_b's stores have completed
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(&flags[b], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
static int diag_inverses(qd_context* h, const T* W, long long ld, long long N, T* inv) {
  constexpr int TB = TriTB<T>::v;
  const unsigned nblk = (unsigned)((N + TB - 1) / TB);
  hipLaunchKernelGGL((k_diag_inverse<T, TB>), dim3(nblk), dim3(TB), 0, h->stream, W, ld, N, 1, inv);
  hipLaunchKernelGGL((k_diag_inverse<T, TB>), dim3(nblk), dim3(TB), 0, h->stream, W, ld, N, 0,
                     inv + (size_t)nblk * TB * TB);
  QD_HIP(h, hipGetLastError());
  return QD_OK;
}

// nrhs <= kBuMaxRhs (lu_solve solves in groups of 8).
// QUADESTH_LOOKBACK=0 keeps the block-step solves (one launch per block).
static bool use_lookback() {
  static const bool on = [] {
    const char* e = std::getenv("QUADESTH_LOOKBACK");
    return !(e && e[0] == '0');
  }();
  return on;
}

template <typename T>
static int tri_solve_inv(qd_context* h, const T* W, long long ld, long long N, int nrhs, T* x, const T* inv) {
  constexpr int TB = TriTB<T>::v;
  if (nrhs < 1 || nrhs > kBuMaxRhs) return set_error(h, QD_ERR_ARG, "tri_solve_inv: nrhs=%d", nrhs);
  const long long nblk = (N + TB - 1) / TB;
  const T* invL = inv;
  const T* invU = inv + (size_t)nblk * TB * TB;
  auto bm = [&](long long b) { return (int)std::min<long long>(TB, N - b * TB); };
  // the solution of the forward sweep goes to tmp and is the backward
  // sweep's working vector; the backward sweep writes its solution to x
  QD_ALLOC(h, h->lu->trs_tmp, (size_t)N * nrhs * sizeof(T));
  T* tmp = reinterpret_cast<T*>(h->lu->trs_tmp.p);
  if (nrhs == 1 && use_lookback()) {
    qd_lu_state* s = h->lu;
    const long long nflags = 1 + nblk * (1 + kLbParts);
    if ((long long)s->lb_flags.count < nflags) {
      QD_ALLOC(h, s->lb_flags, (size_t)nflags);
      QD_HIP(h, hipMemsetAsync(s->lb_flags.p, 0, s->lb_flags.count * sizeof(int), h->stream));
    }
    QD_ALLOC(h, s->lb_part, (size_t)nblk * kLbParts * TB * sizeof(T));
    unsigned long long* pub = nullptr;
    if (sizeof(T) == 4) {
      if ((long long)s->lb_pub.count < nblk * TB) {
        QD_ALLOC(h, s->lb_pub, (size_t)(nblk * TB));
        QD_HIP(h, hipMemsetAsync(s->lb_pub.p, 0, s->lb_pub.count * sizeof(unsigned long long), h->stream));
      }
      pub = s->lb_pub.p;
    }
    T* part = reinterpret_cast<T*>(s->lb_part.p);
    int* err = s->lb_flags.p;  // the epoch of a call in which a wait gave up
    int* flags = s->lb_flags.p + 1;
    int ep = ++s->lb_epoch;
    const dim3 g((unsigned)(nblk * kLbParts));
    hipLaunchKernelGGL((k_trsv_lookback<T, TB>), g, dim3(256), 0, h->stream, W, ld, N, 1, invL, (const T*)x, tmp, part,
                       flags, pub, ep, err);
    ep = ++s->lb_epoch;
    hipLaunchKernelGGL((k_trsv_lookback<T, TB>), g, dim3(256), 0, h->stream, W, ld, N, 0, invU, (const T*)tmp, x, part,
                       flags, pub, ep, err);
    QD_HIP(h, hipGetLastError());
    return QD_OK;
  }
  for (long long b = 0; b < nblk; ++b) {
    const long long i0 = b * TB, r0 = i0 + bm(b);
    const unsigned g = std::max(1u, (unsigned)((N - r0 + kBuRows - 1) / kBuRows));
    hipLaunchKernelGGL((k_block_step<T, TB>), dim3(g), dim3(256), 0, h->stream, W, ld, N, r0, N, i0, bm(b),
                       invL + (size_t)b * TB * TB, nrhs, x, tmp);
  }
  for (long long b = nblk - 1; b >= 0; --b) {
    const long long i0 = b * TB;
    const unsigned g = std::max(1u, (unsigned)((i0 + kBuRows - 1) / kBuRows));
    hipLaunchKernelGGL((k_block_step<T, TB>), dim3(g), dim3(256), 0, h->stream, W, ld, N, 0LL, i0, i0, bm(b),
                       invU + (size_t)b * TB * TB, nrhs, tmp, x);
  }
  QD_HIP(h, hipGetLastError());
  return QD_OK;
}

template <typename T>
static int tri_solve(qd_context* h, const T* W, long long ld, long long N, int nrhs, T* x) {
  // forward, unit lower: block step then update the rows below
  for (long long i0 = 0; i0 < N; i0 += kTB) {
    const int m = (int)std::min<long long>(kTB, N - i0);
    hipLaunchKernelGGL(k_trsv_block<T>, dim3(1), dim3(256), 0, h->stream, W, ld, N, i0, m, 1, nrhs, x);
    if (i0 + m < N)
      hipLaunchKernelGGL(k_trsv_update<T>, dim3((unsigned)((N - i0 - m + 63) / 64)), dim3(256), 0, h->stream, W, ld,
                         N, i0 + m, N, i0, m, nrhs, x);
  }
  // backward, upper
  const long long last = ((N - 1) / kTB) * kTB;
  for (long long i0 = last; i0 >= 0; i0 -= kTB) {
    const int m = (int)std::min<long long>(kTB, N - i0);
    hipLaunchKernelGGL(k_trsv_block<T>, dim3(1), dim3(256), 0, h->stream, W, ld, N, i0, m, 0, nrhs, x);
    if (i0 > 0)
      hipLaunchKernelGGL(k_trsv_update<T>, dim3((unsigned)((i0 + 63) / 64)), dim3(256), 0, h->stream, W, ld, N, 0LL,
                         i0, i0, m, nrhs, x);
  }
  QD_HIP(h, hipGetLastError());
  return QD_OK;
}

// Transposed solve with P A = L U: A^T y = c  <=>  U^T L^T (P y) = c.
// U^T is lower (non-unit), L^T upper (unit).  Column access of row-major
// factors: the update kernel reads W[cols, rows] instead.
template <typename T>
__global__ __launch_bounds__(256) void k_trsvT_block(const T* __restrict__ W, long long ld, long long N, long long i0,
                                                     int m, int upperU, int nrhs, T* __restrict__ x) {
  __shared__ T sL[kTB][kTB + 1];
  __shared__ T sx[8][kTB];
  for (int i = threadIdx.x; i < m * m; i += 256) {
    const int a = i / m, b = i % m;
    sL[b][a] = W[(i0 + a) * ld + i0 + b];  // transposed
  }
  for (int i = threadIdx.x; i < m * nrhs; i += 256) sx[i / m][i % m] = x[(long long)(i / m) * N + i0 + i % m];
  __syncthreads();
  const int r = threadIdx.x % kTB, q = threadIdx.x / kTB;
  if (upperU) {  // U^T: lower, non-unit, forward
    for (int k = 0; k < m; ++k) {
      if (r == k)
        for (int v = q; v < nrhs; v += 4) sx[v][k] = sx[v][k] / sL[k][k];
      __syncthreads();
      if (r > k && r < m)
        for (int v = q; v < nrhs; v += 4) sx[v][r] -= sL[r][k] * sx[v][k];
      __syncthreads();
    }
  } else {  // L^T: upper, unit, backward
    for (int k = m - 1; k >= 0; --k) {
      if (r < k)
        for (int v = q; v < nrhs; v += 4) sx[v][r] -= sL[r][k] * sx[v][k];
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < m * nrhs; i += 256) x[(long long)(i / m) * N + i0 + i % m] = sx[i / m][i % m];
}

// x[rows] -= W[cols, rows]^T x[cols] : rows [r0, r1), cols [c0, c0+m)
template <typename T>
__global__ __launch_bounds__(256) void k_trsvT_update(const T* __restrict__ W, long long ld, long long N,
                                                      long long r0, long long r1, long long c0, int m, int nrhs,
                                                      T* __restrict__ x) {
  __shared__ T sx[8][kTB];
  for (int i = threadIdx.x; i < m * nrhs; i += 256) sx[i / m][i % m] = x[(long long)(i / m) * N + c0 + i % m];
  __syncthreads();
  const long long r = r0 + (long long)blockIdx.x * 256 + threadIdx.x;
  if (r >= r1) return;
  for (int v = 0; v < nrhs; ++v) {
    T s = 0;
    for (int b = 0; b < m; ++b) s += W[(c0 + b) * ld + r] * sx[v][b];
    x[(long long)v * N + r] -= s;
  }
}

template <typename T>
static int tri_solve_T(qd_context* h, const T* W, long long ld, long long N, int nrhs, T* x) {
  for (long long i0 = 0; i0 < N; i0 += kTB) {
    const int m = (int)std::min<long long>(kTB, N - i0);
    hipLaunchKernelGGL(k_trsvT_block<T>, dim3(1), dim3(256), 0, h->stream, W, ld, N, i0, m, 1, nrhs, x);
    if (i0 + m < N)
      hipLaunchKernelGGL(k_trsvT_update<T>, dim3((unsigned)((N - i0 - m + 255) / 256)), dim3(256), 0, h->stream, W,
                         ld, N, i0 + m, N, i0, m, nrhs, x);
  }
  const long long last = ((N - 1) / kTB) * kTB;
  for (long long i0 = last; i0 >= 0; i0 -= kTB) {
    const int m = (int)std::min<long long>(kTB, N - i0);
    hipLaunchKernelGGL(k_trsvT_block<T>, dim3(1), dim3(256), 0, h->stream, W, ld, N, i0, m, 0, nrhs, x);
    if (i0 > 0)
      hipLaunchKernelGGL(k_trsvT_update<T>, dim3((unsigned)((i0 + 255) / 256)), dim3(256), 0, h->stream, W, ld, N,
                         0LL, i0, i0, m, nrhs, x);
  }
  QD_HIP(h, hipGetLastError());
  return QD_OK;
}

// out[v][i] = (T) in[v][perm[i]]   (perm of P: (P b)[i] = b[perm[i]])
template <typename Tin, typename Tout>
__global__ void k_gather(long long N, int nrhs, const int* __restrict__ perm, const Tin* __restrict__ in, long long ldi,
                         Tout* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const long long p = perm[i];
  for (int v = 0; v < nrhs; ++v) out[(long long)v * N + i] = (Tout)in[(long long)v * ldi + p];
}
// out[v][perm[i]] = in[v][i]  (P^T)
template <typename Tin, typename Tout>
__global__ void k_scatter(long long N, int nrhs, const int* __restrict__ perm, const Tin* __restrict__ in,
                          Tout* __restrict__ out, long long ldo) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const long long p = perm[i];
  for (int v = 0; v < nrhs; ++v) out[(long long)v * ldo + p] = (Tout)in[(long long)v * N + i];
}

template <typename Tin, typename Tout>
__global__ void k_convert(long long N, int nrhs, const Tin* __restrict__ in, Tout* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int v = 0; v < nrhs; ++v) out[(long long)v * N + i] = (Tout)in[(long long)v * N + i];
}

// r = b - op(A) x, column-major FP64 A; partial sums over column slices.
// trans = 0: r_i = b_i - sum_c A[i,c] x_c; trans = 1: r_i = b_i - sum_c A[c,i] x_c.
constexpr int kResSlice = 2048;
__global__ __launch_bounds__(256) void k_resid_part(long long N, const double* __restrict__ A, long long lda,
                                                    int nrhs, const double* __restrict__ x, long long ldx,
                                                    double* __restrict__ part) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long c0 = (long long)blockIdx.y * kResSlice;
  const long long c1 = std::min<long long>(N, c0 + kResSlice);
  if (i >= N) return;
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (long long c = c0; c < c1; ++c) {
    const double a = A[c * lda + i];
    for (int v = 0; v < nrhs; ++v) s[v] += a * x[(long long)v * ldx + c];
  }
  for (int v = 0; v < nrhs; ++v) part[((long long)blockIdx.y * nrhs + v) * N + i] = s[v];
}
// One right-hand side: two rows per lane (16-byte loads of an even row pair,
// A 16-byte aligned with an even lda) and eight columns in flight.  Each row
// sums its columns in the same order as k_resid_part.  A is read once per
// sweep, so its loads are non-temporal.
__global__ __launch_bounds__(256) void k_resid1_part(long long N, const double* __restrict__ A, long long lda,
                                                     const double* __restrict__ x, double* __restrict__ part) {
  const long long i = ((long long)blockIdx.x * 256 + threadIdx.x) * 2;
  const long long c0 = (long long)blockIdx.y * kResSlice;
  const long long c1 = std::min<long long>(N, c0 + kResSlice);
  if (i >= N) return;
  const bool pair = i + 1 < N;
  double s0 = 0.0, s1 = 0.0;
  long long c = c0;
  for (; c + 8 <= c1; c += 8) {
    double2 a[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const double* src = A + (c + u) * lda + i;
      if (pair) {
        typedef double d2v __attribute__((ext_vector_type(2)));
        const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v*>(src));
        a[u] = make_double2(v.x, v.y);
      } else {
        a[u] = make_double2(__builtin_nontemporal_load(src), 0.0);
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const double xc = x[c + u];
      s0 += a[u].x * xc;
      s1 += a[u].y * xc;
    }
  }
  for (; c < c1; ++c) {
    const double* src = A + c * lda + i;
    const double xc = x[c];
    s0 += src[0] * xc;
    if (pair) s1 += src[1] * xc;
  }
  part[(long long)blockIdx.y * N + i] = s0;
  if (pair) part[(long long)blockIdx.y * N + i + 1] = s1;
}
// transposed: one wave per output i over a column of A (contiguous)
__global__ __launch_bounds__(256) void k_residT_part(long long N, const double* __restrict__ A, long long lda,
                                                     int nrhs, const double* __restrict__ x, long long ldx,
                                                     double* __restrict__ part) {
  const long long i = (long long)blockIdx.x * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (i >= N) return;
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (long long c = lane; c < N; c += 64) {
    const double a = A[i * lda + c];
    for (int v = 0; v < nrhs; ++v) s[v] += a * x[(long long)v * ldx + c];
  }
  for (int v = 0; v < nrhs; ++v) {
    double t = s[v];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    if (lane == 0) part[(long long)v * N + i] = t;
  }
}
__global__ void k_resid_final(long long N, int nslice, int nrhs, const double* __restrict__ part,
                              const double* __restrict__ b, long long ldb, double* __restrict__ r) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int v = 0; v < nrhs; ++v) {
    double s = 0.0;
    for (int k = 0; k < nslice; ++k) s += part[((long long)k * nrhs + v) * N + i];
    r[(long long)v * N + i] = b[(long long)v * ldb + i] - s;
  }
}
// row sums of |A| (column-major) over column slices: ||A||_inf
__global__ __launch_bounds__(256) void k_rowabs_part(long long N, const double* __restrict__ A, long long lda,
                                                     double* __restrict__ part) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long c0 = (long long)blockIdx.y * kResSlice;
  const long long c1 = std::min<long long>(N, c0 + kResSlice);
  if (i >= N) return;
  double s = 0.0;
  for (long long c = c0; c < c1; ++c) s += fabs(A[c * lda + i]);
  part[(long long)blockIdx.y * N + i] = s;
}

// ---------------------------------------------------------------------------
static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int ensure_factors(qd_context* h, int precision, qd_solve_report* rep) {
  if (!h->lu) h->lu = new qd_lu_state();
  qd_lu_state* s = h->lu;
  if (s->valid && s->precision == precision) return QD_OK;
  const long long N = 3LL * h->n;
  const long long ld = (N + 63) / 64 * 64;
  s->N = N;
  s->ld = ld;
  s->valid = false;
  s->anorm = -1.0;
  s->rowabs_nsl = 0;
  const double t0 = now_ms();
  dim3 tg((unsigned)((N + 31) / 32), (unsigned)((N + 31) / 32));
  int rc;
  // Panel width: the FP16 trailing update streams its FP32 C once per panel,
  // so it takes a wider panel than the MFMA f64/f32 updates need.  With the
  // look-ahead, measured on MI355X at 76,806 DOF: 4096 columns while more
  // than 30,000 remain (the update hides the panel there), 1024 after.  With
  // the wide-tile update, 4096 gave a factor 3-5 ms shorter than 2048 on two
  // boxes (3072: 8 ms longer); 2048 had been 8% shorter than 1024 throughout.
  PanelWidths pw;
  pw.nb = precision == QD_FP16_IR ? 1024 : 512;
  if (precision == QD_FP16_IR) {
    pw.wide = 4096;
    pw.wide_rem = 30000;
  }
  if (const char* e = std::getenv("QUADESTH_NB")) pw.nb = std::max(64, std::atoi(e) / 32 * 32);
  if (precision == QD_FP16_IR) {
    if (const char* e = std::getenv("QUADESTH_NB_WIDE")) pw.wide = std::max(0, std::atoi(e) / 32 * 32);
    if (const char* e = std::getenv("QUADESTH_NB_WIDE_REM")) pw.wide_rem = std::max(0LL, std::atoll(e));
    if (const char* e = std::getenv("QUADESTH_NB_TAIL")) pw.tail = std::max(0, std::atoi(e) / 32 * 32);
    if (const char* e = std::getenv("QUADESTH_NB_TAIL_REM")) pw.tail_rem = std::max(0LL, std::atoll(e));
  }
  if (precision == QD_FP64) {
    s->W32.release();
    QD_ALLOC(h, s->W64, (size_t)ld * (size_t)N);
    hipLaunchKernelGGL(k_transpose_in<double>, tg, dim3(256), 0, h->stream, N, h->A.p, h->lda, s->W64.p, ld);
    QD_HIP(h, hipGetLastError());
    LuCtx<double> c{h, s, s->W64.p, ld, N, false, false, h->stream, &s->sc_main};
    rc = factor(c, pw);
  } else {
    s->W64.release();
    QD_ALLOC(h, s->W32, (size_t)ld * (size_t)N);
    // Diagonal pivots first: the factors only precondition the FP64
    // refinement, and the leaves then skip the pivot search.  The largest
    // multiplier decides afterwards; above kDiagMaxMult (or on a zero pivot)
    // the factorisation is redone with tournament pivoting.
    bool diag = diag_pivots_allowed();
    for (int attempt = 0; attempt < 2; ++attempt) {
      // factor_steps converts A into W32, beside the first panel
      LuCtx<float> c{h, s, s->W32.p, ld, N, precision == QD_FP16_IR, true, h->stream, &s->sc_main};
      c.diag = diag;
      // QD_ASSEMBLY_FAR_FP32_FACTOR: the factors come from the FP32 far-field
      // matrix; the refinement's residual keeps the FP64 A
      c.src = h->have_pre ? h->A_pre.p : h->A.p;
      c.src_ld = h->lda;
      rc = factor(c, pw);
      if (rc || !diag) break;
      unsigned gb = 0;
      QD_HIP(h, hipMemcpyAsync(&gb, s->growth.p, sizeof(gb), hipMemcpyDeviceToHost, h->stream));
      QD_HIP(h, hipStreamSynchronize(h->stream));
      float g;
      std::memcpy(&g, &gb, sizeof(g));
      if (std::getenv("QUADESTH_DEBUG_PIVOT")) std::fprintf(stderr, "[quadesth] diagonal pivots: max |l| = %g\n", g);
      if (g <= kDiagMaxMult) break;  // NaN and +inf fail the test too
      diag = false;
    }
    s->diag_pivots = diag;
    // the conversion's row sums are of the matrix it converted: ||A||_inf from
    // them only when that is A itself
    if (!rc && !h->have_pre) s->rowabs_nsl = (int)(((N + 31) / 32 + kAbsTiles - 1) / kAbsTiles);
  }
  if (rc) return rc;
  {
    if (precision == QD_FP64) {
      constexpr int TB = TriTB<double>::v;
      QD_ALLOC(h, s->inv64, (size_t)2 * ((N + TB - 1) / TB) * TB * TB);
      if ((rc = diag_inverses<double>(h, s->W64.p, ld, N, s->inv64.p))) return rc;
    } else {
      constexpr int TB = TriTB<float>::v;
      QD_ALLOC(h, s->inv32, (size_t)2 * ((N + TB - 1) / TB) * TB * TB);
      if ((rc = diag_inverses<float>(h, s->W32.p, ld, N, s->inv32.p))) return rc;
    }
  }
  // interchanges -> permutation
  s->ipiv.resize(N);
  int st[2];
  QD_HIP(h, hipMemcpyAsync(s->ipiv.data(), s->ipiv_dev.p, N * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  QD_HIP(h, hipMemcpyAsync(st, s->piv_state.p, sizeof(st), hipMemcpyDeviceToHost, h->stream));
  QD_HIP(h, hipStreamSynchronize(h->stream));
  s->info = st[1];
  std::vector<int> perm(N);
  for (long long i = 0; i < N; ++i) perm[i] = (int)i;
  for (long long i = 0; i < N; ++i) std::swap(perm[i], perm[s->ipiv[i]]);
  QD_ALLOC(h, s->perm_dev, (size_t)N);
  QD_HIP(h, hipMemcpy(s->perm_dev.p, perm.data(), N * sizeof(int), hipMemcpyHostToDevice));
  s->precision = precision;
  s->valid = true;
  if (rep) rep->t_factor_ms = now_ms() - t0;
  return QD_OK;
}

static int residual(qd_context* h, bool trans, int nrhs, const double* x, long long ldx, const double* b,
                    long long ldb, double* r) {
  qd_lu_state* s = h->lu;
  const long long N = s->N;
  const int nslice = trans ? 1 : (int)((N + kResSlice - 1) / kResSlice);
  QD_ALLOC(h, s->rowpart, (size_t)nslice * nrhs * N);
  if (!trans && nrhs == 1 && h->lda % 2 == 0 && (reinterpret_cast<uintptr_t>(h->A.p) & 15) == 0)
    hipLaunchKernelGGL(k_resid1_part, dim3((unsigned)((N + 511) / 512), nslice), dim3(256), 0, h->stream, N, h->A.p,
                       h->lda, x, s->rowpart.p);
  else if (!trans)
    hipLaunchKernelGGL(k_resid_part, dim3((unsigned)((N + 255) / 256), nslice), dim3(256), 0, h->stream, N, h->A.p,
                       h->lda, nrhs, x, ldx, s->rowpart.p);
  else
    hipLaunchKernelGGL(k_residT_part, dim3((unsigned)((N + 3) / 4)), dim3(256), 0, h->stream, N, h->A.p, h->lda,
                       nrhs, x, ldx, s->rowpart.p);
  hipLaunchKernelGGL(k_resid_final, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, h->stream, N, nslice, nrhs,
                     s->rowpart.p, b, ldb, r);
  QD_HIP(h, hipGetLastError());
  return QD_OK;
}

// d = op(A)^-1 r with the held factors; r, d: nrhs vectors of stride N (FP64);
// t64: nrhs * N doubles of work space for the FP64 factors.
static int apply_inverse(qd_context* h, bool trans, int nrhs, const double* r, double* d, double* t64) {
  qd_lu_state* s = h->lu;
  const long long N = s->N;
  const unsigned nb = (unsigned)((N + 255) / 256);
  if (s->precision == QD_FP64) {
    double* t = t64;
    if (!trans) {
      hipLaunchKernelGGL((k_gather<double, double>), dim3(nb), dim3(256), 0, h->stream, N, nrhs, s->perm_dev.p, r, N, t);
      int rc = tri_solve_inv<double>(h, s->W64.p, s->ld, N, nrhs, t, s->inv64.p);
      if (rc) return rc;
      QD_HIP(h, hipMemcpyAsync(d, t, (size_t)nrhs * N * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    } else {
      QD_HIP(h, hipMemcpyAsync(t, r, (size_t)nrhs * N * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
      int rc = tri_solve_T<double>(h, s->W64.p, s->ld, N, nrhs, t);
      if (rc) return rc;
      hipLaunchKernelGGL((k_scatter<double, double>), dim3(nb), dim3(256), 0, h->stream, N, nrhs, s->perm_dev.p, t, d, N);
    }
  } else {
    float* t = s->vecf.p;
    if (!trans) {
      hipLaunchKernelGGL((k_gather<double, float>), dim3(nb), dim3(256), 0, h->stream, N, nrhs, s->perm_dev.p, r, N, t);
      int rc = tri_solve_inv<float>(h, s->W32.p, s->ld, N, nrhs, t, s->inv32.p);
      if (rc) return rc;
      hipLaunchKernelGGL((k_convert<float, double>), dim3(nb), dim3(256), 0, h->stream, N, nrhs, t, d);
    } else {
      hipLaunchKernelGGL((k_convert<double, float>), dim3(nb), dim3(256), 0, h->stream, N, nrhs, r, t);
      int rc = tri_solve_T<float>(h, s->W32.p, s->ld, N, nrhs, t);
      if (rc) return rc;
      hipLaunchKernelGGL((k_scatter<float, double>), dim3(nb), dim3(256), 0, h->stream, N, nrhs, s->perm_dev.p, t, d, N);
    }
  }
  QD_HIP(h, hipGetLastError());
  return QD_OK;
}

__global__ void k_axpy_inplace(long long n, const double* __restrict__ d, double* __restrict__ x) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += d[i];
}

__global__ void k_fill_zero(long long n, double* __restrict__ x) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

// max(a, b) that keeps a NaN in either operand
__host__ __device__ __forceinline__ double max_keep_nan(double a, double b) { return (a != a || a > b) ? a : b; }

// Block maxima of |v| (NaN kept) for nrhs vectors of stride ld: part[v * kNormBlocks + block].
constexpr int kNormBlocks = 64;
__global__ __launch_bounds__(256) void k_absmax_part(long long N, const double* __restrict__ v, long long ld,
                                                     double* __restrict__ part) {
  __shared__ double red[256];
  const double* x = v + (long long)blockIdx.y * ld;
  double m = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < N; i += (long long)kNormBlocks * 256)
    m = max_keep_nan(fabs(x[i]), m);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] = max_keep_nan(red[threadIdx.x + k], red[threadIdx.x]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long long)blockIdx.y * kNormBlocks + blockIdx.x] = red[0];
}

// inf-norms of nrhs (<= 8) vectors: block maxima on the device, the last
// kNormBlocks per vector on the host
static int host_norms(qd_context* h, int nrhs, long long N, const double* dev, long long ld, double* out) {
  qd_lu_state* s = h->lu;
  if (nrhs < 1 || nrhs > 8) return set_error(h, QD_ERR_ARG, "host_norms: nrhs=%d", nrhs);
  QD_ALLOC(h, s->nrm, (size_t)8 * kNormBlocks);
  hipLaunchKernelGGL(k_absmax_part, dim3(kNormBlocks, nrhs), dim3(256), 0, h->stream, N, dev, ld, s->nrm.p);
  QD_HIP(h, hipGetLastError());
  double part[8 * kNormBlocks];
  QD_HIP(h, hipMemcpyAsync(part, s->nrm.p, (size_t)nrhs * kNormBlocks * sizeof(double), hipMemcpyDeviceToHost,
                           h->stream));
  QD_HIP(h, hipStreamSynchronize(h->stream));
  for (int v = 0; v < nrhs; ++v) {
    double m = 0.0;
    for (int i = 0; i < kNormBlocks; ++i) m = max_keep_nan(part[v * kNormBlocks + i], m);
    out[v] = m;
  }
  return QD_OK;
}

static int gmres_ir(qd_context* h, bool trans, double* x, double* b, double* t64, double tol, int max_iter,
                    int* iters, double* relres);

// ||A||_inf of the FP64 matrix, once per held factorisation: from the row
// sums that A's conversion into W32 left, else one pass over A.
static int ensure_anorm(qd_context* h) {
  qd_lu_state* s = h->lu;
  const long long N = s->N;
  if (s->anorm >= 0.0) return QD_OK;
  int rc;
  double nn = 0.0;
  if (s->rowabs_nsl > 0) {
    hipLaunchKernelGGL(k_rowabs_sum, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, h->stream, N, s->rowabs_nsl,
                       s->rowabs.p, s->rowabs.p + (size_t)s->rowabs_nsl * N);
    QD_HIP(h, hipGetLastError());
    if ((rc = host_norms(h, 1, N, s->rowabs.p + (size_t)s->rowabs_nsl * N, N, &nn))) return rc;
  } else {
    // slice partials of the row sums of |A|, then their sums (in the residual
    // partials' buffer)
    const int nsl = (int)((N + kResSlice - 1) / kResSlice);
    QD_ALLOC(h, s->rowpart, (size_t)(nsl + 1) * N);
    double* ra = s->rowpart.p;
    hipLaunchKernelGGL(k_rowabs_part, dim3((unsigned)((N + 255) / 256), nsl), dim3(256), 0, h->stream, N, h->A.p,
                       h->lda, ra);
    hipLaunchKernelGGL(k_rowabs_sum, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, h->stream, N, nsl, ra,
                       ra + (size_t)nsl * N);
    QD_HIP(h, hipGetLastError());
    if ((rc = host_norms(h, 1, N, ra + (size_t)nsl * N, N, &nn))) return rc;
    // For A^T the bound would be ||A||_1; the row norm is kept for both (it
    // only scales the stopping test, and the two are within a factor N).
  }
  s->anorm = nn;
  return QD_OK;
}

// Largest normwise backward error over the right-hand sides:
// ||r_v||_inf / (||A||_inf ||x_v||_inf + ||b_v||_inf); NaN anywhere gives NaN.
static int backward_error(qd_context* h, int nrhs, const double* Rr, const double* X, const double* bn,
                          double* worst) {
  qd_lu_state* s = h->lu;
  const long long N = s->N;
  double rn[8], xn[8];
  int rc;
  if ((rc = host_norms(h, nrhs, N, Rr, N, rn))) return rc;
  if ((rc = host_norms(h, nrhs, N, X, N, xn))) return rc;
  double w = 0.0;
  for (int v = 0; v < nrhs; ++v) w = max_keep_nan(rn[v] / (s->anorm * xn[v] + bn[v] + 1e-300), w);
  *worst = w;
  return QD_OK;
}

// The refinement's default bound on the normwise backward error
// ||b - Ax||inf / (||A||inf ||x||inf + ||b||inf): LAPACK's mixed-precision
// DSGESV (dsgesv.f) stops once ||r||inf < ||x||inf ||A||inf eps sqrt(n), eps =
// 2^-53; with ||b|| in its denominator this bound is at least as strict.
// Below 8,100 unknowns it is 1e-14.
static double default_ir_tol(long long N) { return std::max(1e-14, std::ldexp(1.0, -53) * std::sqrt((double)N)); }

int lu_solve(qd_context* h, int precision, int nrhs, double* b, long long ldb, double tol, int max_iter,
             bool trans, qd_solve_report* rep) {
  if (nrhs > 8) {
    // rhs blocks of 8
    int rc = QD_OK;
    qd_solve_report worst{};
    for (int v0 = 0; v0 < nrhs; v0 += 8) {
      qd_solve_report r{};
      rc = lu_solve(h, precision, std::min(8, nrhs - v0), b + (long long)v0 * ldb, ldb, tol, max_iter, trans, &r);
      if (rc < 0) return rc;
      worst.precision = r.precision;
      worst.info = std::max(worst.info, r.info);
      worst.iterations = std::max(worst.iterations, r.iterations);
      worst.converged = (v0 == 0) ? r.converged : (worst.converged && r.converged);
      worst.gmres |= r.gmres;
      worst.pivoting |= r.pivoting;
      worst.residual = std::max(worst.residual, r.residual);
      worst.t_factor_ms += r.t_factor_ms;
      worst.t_solve_ms += r.t_solve_ms;
    }
    if (rep) *rep = worst;
    return rc;
  }
  qd_solve_report local{};
  qd_solve_report* R = rep ? rep : &local;
  *R = qd_solve_report{};
  R->precision = precision;
  int rc = ensure_factors(h, precision, R);
  if (rc) return rc;
  qd_lu_state* s = h->lu;
  R->info = s->info;
  R->pivoting = (precision != QD_FP64 && s->diag_pivots) ? 1 : 0;
  if (s->info > 0) return s->info;
  const long long N = s->N;
  const double t0 = now_ms();
  const int lb_epoch0 = s->lb_epoch;
  const long long S = (long long)nrhs * N;
  QD_ALLOC(h, s->vec, (size_t)5 * S);
  QD_ALLOC(h, s->vecf, (size_t)S);
  double* X = s->vec.p;                 // solutions
  double* Rr = s->vec.p + S;            // residuals
  double* D = s->vec.p + 2 * S;         // corrections
  double* Bc = s->vec.p + 3 * S;        // compact copy of b
  double* T64 = s->vec.p + 4 * S;       // work vectors of the FP64 substitutions
  for (int v = 0; v < nrhs; ++v)
    QD_HIP(h, hipMemcpyAsync(Bc + (long long)v * N, b + (long long)v * ldb, N * sizeof(double),
                             hipMemcpyDeviceToDevice, h->stream));
  if ((rc = ensure_anorm(h))) return rc;
  double bn[8];
  if ((rc = host_norms(h, nrhs, N, Bc, N, bn))) return rc;
  if (precision == QD_FP64) {
    rc = apply_inverse(h, trans, nrhs, Bc, X, T64);
    if (rc) return rc;
    // the direct solve's own backward error against the FP64 A (one residual)
    if ((rc = residual(h, trans, nrhs, X, N, Bc, N, Rr))) return rc;
    if ((rc = backward_error(h, nrhs, Rr, X, bn, &R->residual))) return rc;
    R->converged = 1;
    R->iterations = 0;
  } else {
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, h->stream, S, X);
    QD_HIP(h, hipMemcpyAsync(Rr, Bc, (size_t)S * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    double prev = 1e300;
    int it = 0;
    bool done = false, stalled = false;
    const int maxit = max_iter > 0 ? max_iter : 30;
    const double tl = tol > 0.0 ? tol : default_ir_tol(N);
    // QUADESTH_IR_SWEEPS=k hands the solve to GMRES after k plain sweeps
    // that have not met the tolerance (the fallback's own trigger is a sweep
    // that fails to halve the residual)
    int ir_sweeps = maxit;
    if (const char* e = std::getenv("QUADESTH_IR_SWEEPS")) ir_sweeps = std::max(1, std::min(maxit, std::atoi(e)));
    while (it < maxit) {
      if (it == ir_sweeps) {
        stalled = true;
        break;
      }
      rc = apply_inverse(h, trans, nrhs, Rr, D, T64);
      if (rc) return rc;
      hipLaunchKernelGGL(k_axpy_inplace, dim3((unsigned)((nrhs * N + 255) / 256)), dim3(256), 0, h->stream,
                         (long long)nrhs * N, D, X);
      ++it;
      rc = residual(h, trans, nrhs, X, N, Bc, N, Rr);
      if (rc) return rc;
      double worst = 0.0;
      if ((rc = backward_error(h, nrhs, Rr, X, bn, &worst))) return rc;
      R->residual = worst;
      if (worst <= tl) {
        done = true;
        break;
      }
      if (!(worst < 0.5 * prev)) {
        // no progress, or a non-finite residual (GMRES stops on that too)
        stalled = true;
        break;
      }
      prev = worst;
    }
    R->iterations = it;
    R->converged = done ? 1 : 0;
    if (!done && stalled) {
      // GMRES fallback per rhs, started from the current iterate
      R->gmres = 1;
      double worst = 0.0;
      bool all = true;
      for (int v = 0; v < nrhs; ++v) {
        int its = 0;
        double rr = 0.0;
        rc = gmres_ir(h, trans, X + (long long)v * N, Bc + (long long)v * N, T64, tl, maxit, &its, &rr);
        if (rc < 0) return rc;
        all = all && rr <= tl;
        worst = std::max(worst, rr);
        R->iterations += its;
      }
      R->residual = worst;
      R->converged = all ? 1 : 0;
    }
  }
  for (int v = 0; v < nrhs; ++v)
    QD_HIP(h, hipMemcpyAsync(b + (long long)v * ldb, X + (long long)v * N, N * sizeof(double),
                             hipMemcpyDeviceToDevice, h->stream));
  QD_HIP(h, hipStreamSynchronize(h->stream));
  R->t_solve_ms = now_ms() - t0;
  if (s->lb_epoch != lb_epoch0) {
    // a look-back solve whose wait gave up leaves its epoch in the error word
    int lb_err = 0;
    QD_HIP(h, hipMemcpy(&lb_err, s->lb_flags.p, sizeof(int), hipMemcpyDeviceToHost));
    if (lb_err > lb_epoch0)
      return set_error(h, QD_ERR_HIP, "look-back triangular solve: a wait for an earlier block timed out (call %d)",
                       lb_err);
  }
  if (precision != QD_FP64 && !R->converged)
    return set_error(h, QD_ERR_NOCONV, "refinement stopped at relative residual %.3e (tol %.1e) after %d sweeps",
                     R->residual, tol > 0 ? tol : default_ir_tol(N), R->iterations);
  return QD_OK;
}

// ---------------------------------------------------------------------------
// Restarted GMRES on M^-1 op(A) x = M^-1 b (left preconditioning with the
// low-precision factors), FP64 Arnoldi with modified Gram-Schmidt.  Dot
// products and norms go through small device reductions.
__global__ __launch_bounds__(256) void k_dot_part(long long n, const double* __restrict__ a,
                                                  const double* __restrict__ b, double* __restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) s += a[i] * b[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void k_axpby(long long n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

// *out = a . b; an error code when the device reduction fails, QD_ERR_NOCONV
// when the value is not finite (the Arnoldi process cannot continue).
static int dot(qd_context* h, long long n, const double* a, const double* b, double* out) {
  const int nb = 512;
  DevBuf<double>& part = h->lu->rowpart;
  QD_ALLOC(h, part, std::max<size_t>(part.count, nb));
  hipLaunchKernelGGL(k_dot_part, dim3(nb), dim3(256), 0, h->stream, n, a, b, part.p);
  QD_HIP(h, hipGetLastError());
  double hp[nb];
  QD_HIP(h, hipMemcpyAsync(hp, part.p, sizeof(hp), hipMemcpyDeviceToHost, h->stream));
  QD_HIP(h, hipStreamSynchronize(h->stream));
  double s = 0.0;
  for (int i = 0; i < nb; ++i) s += hp[i];
  if (!std::isfinite(s)) return set_error(h, QD_ERR_NOCONV, "GMRES: non-finite inner product");
  *out = s;
  return QD_OK;
}

constexpr int kGmresM = 20;  // Krylov vectors per restart
static int gmres_ir(qd_context* h, bool trans, double* x, double* b, double* t64, double tol, int max_restart,
                    int* iters, double* relres) {
  qd_lu_state* s = h->lu;
  const long long N = s->N;
  constexpr int m = kGmresM;
  QD_ALLOC(h, s->gm, (size_t)(m + 3) * N);
  double* V = s->gm.p;
  double* w = V + (size_t)(m + 1) * N;
  double* r = V + (size_t)(m + 2) * N;
  const unsigned nb = (unsigned)((N + 255) / 256);
  hipLaunchKernelGGL(k_fill_zero, dim3(nb), dim3(256), 0, h->stream, N, V + (size_t)m * N);
  double bn[1];
  int rc;
  if ((rc = host_norms(h, 1, N, b, N, bn))) return rc;
  *iters = 0;
  for (int restart = 0; restart < std::max(1, max_restart); ++restart) {
    if ((rc = residual(h, trans, 1, x, N, b, N, r))) return rc;
    if ((rc = backward_error(h, 1, r, x, bn, relres))) return rc;
    if (*relres <= tol) return QD_OK;
    if (!std::isfinite(*relres)) return set_error(h, QD_ERR_NOCONV, "GMRES: non-finite residual");
    rc = apply_inverse(h, trans, 1, r, w, t64);
    if (rc) return rc;
    double ww;
    if ((rc = dot(h, N, w, w, &ww))) return rc;
    const double beta = std::sqrt(ww);
    double H[(m + 1) * m] = {}, cs[m] = {}, sn[m] = {}, g[m + 1] = {};
    g[0] = beta;
    hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(256), 0, h->stream, N, 1.0 / beta, w, 0.0, V);
    int k = 0;
    for (; k < m; ++k) {
      // w = M^-1 A v_k
      rc = residual(h, trans, 1, V + (size_t)k * N, N, V + (size_t)m * N /*dummy b*/, N, r);
      if (rc) return rc;
      // residual gives b - A v; use b = 0 slot: compute A v = -(0 - A v)
      // (V[m] is zeroed below before first use)
      rc = apply_inverse(h, trans, 1, r, w, t64);
      if (rc) return rc;
      hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(256), 0, h->stream, N, 0.0, w, -1.0, w);
      for (int j = 0; j <= k; ++j) {
        double hij;
        if ((rc = dot(h, N, w, V + (size_t)j * N, &hij))) return rc;
        H[(size_t)j * m + k] = hij;
        hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(256), 0, h->stream, N, -hij, V + (size_t)j * N, 1.0, w);
      }
      double hh;
      if ((rc = dot(h, N, w, w, &hh))) return rc;
      const double hn = std::sqrt(hh);
      H[(size_t)(k + 1) * m + k] = hn;
      if (k + 1 < m) hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(256), 0, h->stream, N, 1.0 / hn, w, 0.0, V + (size_t)(k + 1) * N);
      for (int j = 0; j < k; ++j) {
        const double t = cs[j] * H[(size_t)j * m + k] + sn[j] * H[(size_t)(j + 1) * m + k];
        H[(size_t)(j + 1) * m + k] = -sn[j] * H[(size_t)j * m + k] + cs[j] * H[(size_t)(j + 1) * m + k];
        H[(size_t)j * m + k] = t;
      }
      const double a = H[(size_t)k * m + k], bb = H[(size_t)(k + 1) * m + k];
      const double rr = std::sqrt(a * a + bb * bb);
      cs[k] = a / rr;
      sn[k] = bb / rr;
      H[(size_t)k * m + k] = rr;
      H[(size_t)(k + 1) * m + k] = 0.0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      ++*iters;
      if (std::fabs(g[k + 1]) <= 1e-3 * tol * beta || k + 1 == m) {
        ++k;
        break;
      }
    }
    double y[m] = {};
    for (int i = k - 1; i >= 0; --i) {
      double t = g[i];
      for (int j = i + 1; j < k; ++j) t -= H[(size_t)i * m + j] * y[j];
      y[i] = t / H[(size_t)i * m + i];
    }
    for (int i = 0; i < k; ++i)
      hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(256), 0, h->stream, N, y[i], V + (size_t)i * N, 1.0, x);
    QD_HIP(h, hipGetLastError());
  }
  if ((rc = residual(h, trans, 1, x, N, b, N, r))) return rc;
  return backward_error(h, 1, r, x, bn, relres);
}

void lu_free(qd_context* h) {
  delete h->lu;
  h->lu = nullptr;
}

// ---------------------------------------------------------------------------
// Allocation-free solves (qd_solve_device_noalloc).  Every work vector of a
// solve with held factors comes from this one static array: each GPU the
// library runs on holds its own copy of it until the process ends.  Host code
// reaches a GPU's copy only through the address hipGetSymbolAddress gave for
// it (the array's name is a host-side stand-in, not a device address).  The
// array keeps the last call's contents.  Its first part holds only the
// look-back solves' flags (with the error word) and published words, at the
// same offsets in every call: they carry a per-device epoch that only grows,
// so a stale word never matches a later call; the array is cleared once at
// set-up.  The work vectors follow, sliced per call by N and the number of
// right-hand sides.  Calls that overlap on one GPU share the array and race.
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr size_t kScratchFlags = 16384;    // ints: error word, block and part flags
constexpr size_t kScratchPub = 131072;     // 64-bit published words of the FP32 look-back solves
constexpr size_t kScratchVectors = kScratchFlags * 4 + kScratchPub * 8;  // offset of the work vectors
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[kMaxDevices];
static int g_scratch_epoch[kMaxDevices];

__global__ void k_fill_zero_u32(long long n, unsigned* __restrict__ x) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] = 0u;
}

int noalloc_init(int device, std::string* err) {
  if (device < 0 || device >= kMaxDevices) {
    *err = "device index beyond the allocation-free workspace table";
    return QD_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return QD_OK;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return QD_ERR_HIP;
  }
  hipLaunchKernelGGL(k_fill_zero_u32, dim3(1024), dim3(256), 0, nullptr, (long long)(kScratchBytes / 4),
                     static_cast<unsigned*>(ptr));
  if ((e = hipGetLastError()) == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) {
    *err = std::string("clearing the allocation-free workspace: ") + hipGetErrorString(e);
    return QD_ERR_HIP;
  }
  g_scratch_ptr[device] = static_cast<unsigned char*>(ptr);
  g_scratch_epoch[device] = 0;
  return QD_OK;
}

namespace {
// Points a handle buffer at a slice of the workspace for one call, restoring
// the buffer afterwards.
template <typename T>
struct ViewIn {
  DevBuf<T>& b;
  T* p0;
  size_t c0;
  bool v0;
  ViewIn(DevBuf<T>& buf, unsigned char* ptr, size_t cnt) : b(buf), p0(buf.p), c0(buf.count), v0(buf.view) {
    b.p = reinterpret_cast<T*>(ptr);
    b.count = cnt;
    b.view = true;
  }
  ~ViewIn() {
    b.p = p0;
    b.count = c0;
    b.view = v0;
  }
};
struct ForbidAlloc {
  qd_context* h;
  explicit ForbidAlloc(qd_context* hh) : h(hh) { h->forbid_alloc = true; }
  ~ForbidAlloc() { h->forbid_alloc = false; }
};
}  // namespace

int lu_solve_noalloc(qd_context* h, int precision, int nrhs, double* b, long long ldb, double tol, int max_iter,
                     qd_solve_report* rep) {
  qd_lu_state* s = h->lu;
  if (!s || !s->valid || s->precision != precision)
    return set_error(h, QD_ERR_STATE,
                     "qd_solve_device_noalloc needs the factors of an earlier qd_solve of precision %d on this handle",
                     precision);
  if (h->device < 0 || h->device >= kMaxDevices) return set_error(h, QD_ERR_ARG, "device %d", h->device);
  unsigned char* base;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    base = g_scratch_ptr[h->device];
  }
  if (!base) return set_error(h, QD_ERR_STATE, "allocation-free workspace not set up on device %d", h->device);
  // the slices one call of lu_solve (at most 8 right-hand sides at a time)
  // asks for: QD_ALLOC finds each big enough and allocates nothing
  const long long N = s->N;
  const long long r = std::min(nrhs, 8);
  const bool f64 = precision == QD_FP64;
  const long long TB = f64 ? TriTB<double>::v : TriTB<float>::v;
  const long long esz = f64 ? 8 : 4;
  const long long nblk = (N + TB - 1) / TB;
  const long long nslice = (N + kResSlice - 1) / kResSlice;
  const size_t n_vec = (size_t)(5 * r * N), n_vecf = (size_t)(r * N);
  const size_t n_rowpart = (size_t)std::max<long long>((nslice * r + 1) * N, 512);
  const size_t n_nrm = (size_t)8 * kNormBlocks, n_trs = (size_t)(N * r * 8);
  const size_t n_flags = (size_t)(1 + nblk * (1 + kLbParts)), n_part = (size_t)(nblk * kLbParts * TB * esz);
  const size_t n_pub = f64 ? 0 : (size_t)(nblk * TB), n_gm = (size_t)(kGmresM + 3) * N;
  // flags and published words at fixed offsets; the vectors after them
  unsigned char* p_flags = base;
  unsigned char* p_pub = base + kScratchFlags * 4;
  size_t off = kScratchVectors;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return base + o;
  };
  unsigned char* p_vec = take(n_vec * 8);
  unsigned char* p_vecf = take(n_vecf * 4);
  unsigned char* p_rowpart = take(n_rowpart * 8);
  unsigned char* p_nrm = take(n_nrm * 8);
  unsigned char* p_trs = take(n_trs);
  unsigned char* p_part = take(n_part);
  unsigned char* p_gm = take(n_gm * 8);
  if (off > kScratchBytes || n_flags > kScratchFlags || n_pub > kScratchPub)
    return set_error(h, QD_ERR_ARG,
                     "qd_solve_device_noalloc: N=%lld with %lld right-hand sides at a time needs %zu bytes of "
                     "workspace (%zu flags, %zu published words); the static workspace holds %zu (%zu, %zu)",
                     N, r, off, n_flags, n_pub, kScratchBytes, kScratchFlags, kScratchPub);
  ViewIn<double> v1(s->vec, p_vec, n_vec);
  ViewIn<float> v2(s->vecf, p_vecf, n_vecf);
  ViewIn<double> v3(s->rowpart, p_rowpart, n_rowpart);
  ViewIn<double> v4(s->nrm, p_nrm, n_nrm);
  ViewIn<unsigned char> v5(s->trs_tmp, p_trs, n_trs);
  ViewIn<int> v6(s->lb_flags, p_flags, n_flags);
  ViewIn<unsigned char> v7(s->lb_part, p_part, n_part);
  ViewIn<unsigned long long> v8(s->lb_pub, p_pub, n_pub);
  ViewIn<double> v9(s->gm, p_gm, n_gm);
  ForbidAlloc guard(h);
  // the workspace's flags follow the device's epochs, not the handle's
  const int own_epoch = s->lb_epoch;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    s->lb_epoch = g_scratch_epoch[h->device];
  }
  const int rc = lu_solve(h, precision, nrhs, b, ldb, tol, max_iter, false, rep);
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    g_scratch_epoch[h->device] = std::max(g_scratch_epoch[h->device], s->lb_epoch);
  }
  s->lb_epoch = own_epoch;
  return rc;
}

int fp16_update(qd_context* h, long long M, long long N, int Kp, const void* A16, const void* B16,
                const float* colscale, float* C, long long ldc) {
  if (M <= 0 || N <= 0) return QD_OK;
  const int rows16 = (ldc % 4 == 0 && (reinterpret_cast<uintptr_t>(C) & 15) == 0) ? 1 : 0;
  return launch_f16_update(h, true, M, N, Kp, (const _Float16*)A16, (const _Float16*)B16, colscale, C, ldc, rows16, 0,
                           h->stream);
}

int lu_invalidate(qd_context* h) {
  if (h->lu) {
    h->lu->valid = false;
    h->lu->anorm = -1.0;
  }
  return QD_OK;
}

// DGETRF + DGETRS on caller device memory (column-major), FP64.
int dgesv_device(qd_context* h, int n, int nrhs, double* a, long long lda, int* ipiv_host, double* b, long long ldb) {
  if (!h->lu) h->lu = new qd_lu_state();
  qd_lu_state* s = h->lu;
  s->valid = false;
  const long long N = n;
  const long long ld = (N + 63) / 64 * 64;
  s->N = N;
  s->ld = ld;
  s->W32.release();
  QD_ALLOC(h, s->W64, (size_t)ld * (size_t)N);
  dim3 tg((unsigned)((N + 31) / 32), (unsigned)((N + 31) / 32));
  hipLaunchKernelGGL(k_transpose_in<double>, tg, dim3(256), 0, h->stream, N, a, lda, s->W64.p, ld);
  QD_HIP(h, hipGetLastError());
  LuCtx<double> c{h, s, s->W64.p, ld, N, false, false, h->stream, &s->sc_main};
  PanelWidths pw;
  pw.nb = 512;
  int rc = factor(c, pw);
  if (rc) return rc;
  std::vector<int> ip(N);
  int st[2];
  QD_HIP(h, hipMemcpyAsync(ip.data(), s->ipiv_dev.p, N * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  QD_HIP(h, hipMemcpyAsync(st, s->piv_state.p, sizeof(st), hipMemcpyDeviceToHost, h->stream));
  QD_HIP(h, hipStreamSynchronize(h->stream));
  for (long long i = 0; i < N; ++i) ipiv_host[i] = ip[i] + 1;
  hipLaunchKernelGGL(k_transpose_out, tg, dim3(256), 0, h->stream, N, s->W64.p, ld, a, lda);
  QD_HIP(h, hipGetLastError());
  if (st[1] > 0) {
    QD_HIP(h, hipStreamSynchronize(h->stream));
    return st[1];
  }
  std::vector<int> perm(N);
  for (long long i = 0; i < N; ++i) perm[i] = (int)i;
  for (long long i = 0; i < N; ++i) std::swap(perm[i], perm[ip[i]]);
  QD_ALLOC(h, s->perm_dev, (size_t)N);
  QD_HIP(h, hipMemcpy(s->perm_dev.p, perm.data(), N * sizeof(int), hipMemcpyHostToDevice));
  QD_ALLOC(h, s->vec, (size_t)5 * 8 * N);
  double* t = s->vec.p;
  for (int v0 = 0; v0 < nrhs; v0 += 8) {
    const int nv = std::min(8, nrhs - v0);
    hipLaunchKernelGGL((k_gather<double, double>), dim3((unsigned)((N + 255) / 256)), dim3(256), 0, h->stream, N, nv,
                       s->perm_dev.p, b + (long long)v0 * ldb, ldb, t);
    rc = tri_solve<double>(h, s->W64.p, ld, N, nv, t);
    if (rc) return rc;
    for (int v = 0; v < nv; ++v)
      QD_HIP(h, hipMemcpyAsync(b + (long long)(v0 + v) * ldb, t + (long long)v * N, N * sizeof(double),
                               hipMemcpyDeviceToDevice, h->stream));
  }
  QD_HIP(h, hipStreamSynchronize(h->stream));
  // the handle's factors no longer belong to its assembled system
  s->valid = false;
  return QD_OK;
}

}  // namespace qd
