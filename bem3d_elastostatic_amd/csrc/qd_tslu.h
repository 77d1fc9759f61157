// Tournament pivoting kernels of the FP32 preconditioner LU (a leaf of <= 32
// columns): the selection tree and the finish.  Device code only; included
// by qd_lu.hip inside namespace qd.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>

namespace qd {

// Tournament pivoting for a leaf of <= 32 columns (the factorisation that
// only preconditions the FP64 refinement), in two launches per leaf.
//
// k_tslu_tree runs the whole selection tree.  A node is Gaussian elimination
// with partial pivoting on up to 512 rows, two per lane of four waves, and
// yields its 32 winning row numbers in pivot order.  A step's pivot search is
// a DPP maximum in each wave and one barrier to compare the waves' best rows,
// whose values travel through LDS.  Level 0 has one node per 512 rows of the
// leaf; each level above has one node per 16 nodes below, over their 16 x 32
// winners.  No workgroup waits for another: the last workgroup of
// a group to finish (an agent-scope counter) goes on to run the parent node,
// and the workgroup that runs the root writes the winners' LU, the LAPACK
// interchange sequence (ipiv) and the permutation it amounts to.  The rows are
// read from W, which the tree does not modify, so every node works on the
// original values.
//
// k_tslu_finish then writes the winners' LU into the leaf's top rows, forms
// L21 = A21 U11^-1 for every row below (a row whose content the interchanges
// replace reads that content from the copy of the original top rows the root
// made) and, in its remaining workgroups, permutes the panel's other columns.
//
// Both raise their waves' issue priority (s_setprio): they run beside the
// trailing update on the same SIMDs and sit on the factorisation's critical
// path.
constexpr int kTsW = 32;
// A node of the tree: V waves holding R rows per lane, 64 V R rows; a parent
// takes the winners of fan = rows / 32 children.
template <int R, int V>
struct TsNode {
  static constexpr int threads = 64 * V;
  static constexpr int rows = threads * R;
  static constexpr int fan = rows / kTsW;
};
// Measured on MI355X (one leaf of 76,800 rows: 85 us at 2 x 4 against 98 us
// at 4 x 1): fewer levels outweigh the barrier each step costs with more
// waves, and 120 VGPRs fit beside the trailing update more easily than 228.
#ifndef QD_TS_R
#define QD_TS_R 2
#endif
#ifndef QD_TS_V
#define QD_TS_V 4
#endif
constexpr int kTsTreeR = QD_TS_R, kTsTreeV = QD_TS_V;
static_assert(kTsTreeR >= 1 && kTsTreeR <= 4 && kTsTreeV >= 1 && kTsTreeV <= 16, "tree node shape");
// meta layout: [0, 32) ipiv values, [32, 64) displaced positions, [64, 96)
// the original top row each of them receives, [96] their count, [128, 160)
// the original row that ends at top position c0 + k.
constexpr int kTsMeta = 160;

// Candidate-list entries and counters the tree needs for a leaf of m rows.
template <int R = kTsTreeR, int V = kTsTreeV>
static inline void ts_tree_sizes(long long m, long long& cand, long long& counters) {
  using Nd = TsNode<R, V>;
  long long nl = (m + Nd::rows - 1) / Nd::rows;
  cand = 0;
  counters = 0;
  while (true) {
    cand += nl * kTsW;
    if (nl == 1) break;
    const long long up = (nl + Nd::fan - 1) / Nd::fan;
    counters += up;
    nl = up;
  }
}

// Wave maximum of an unsigned value: DPP butterflies inside each row of 16
// lanes (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror,
// row_mirror), then the four row maxima through readlane.  Called by all 64
// lanes; the result is uniform.
template <int kCtrl>
__device__ __forceinline__ unsigned dpp_max_u32(unsigned v) {
  const unsigned o = (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, kCtrl, 0xf, 0xf, false);
  return o > v ? o : v;
}

__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
  v = dpp_max_u32<0xB1>(v);
  v = dpp_max_u32<0x4E>(v);
  v = dpp_max_u32<0x141>(v);
  v = dpp_max_u32<0x140>(v);
  unsigned m = (unsigned)__builtin_amdgcn_readlane((int)v, 0);
  m = max(m, (unsigned)__builtin_amdgcn_readlane((int)v, 16));
  m = max(m, (unsigned)__builtin_amdgcn_readlane((int)v, 32));
  m = max(m, (unsigned)__builtin_amdgcn_readlane((int)v, 48));
  return m;
}

__device__ __forceinline__ float readlane_f(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// LDS of a node's elimination: per step parity, each wave's candidate (its
// key, global row and the row itself).
template <int V>
struct TsNodeSmem {
  float row[2][V][kTsW];
  unsigned key[2][V];
  int idx[2][V];
};

// A wave's pivot candidate for column K: a row's key is bits(|a|) + 1 (0 = no
// row), so the largest |a| wins; ties go to the lowest lane, then the lane's
// first row.  wmax is the wave's best key, held by row qw of lane wl.
struct TsPick {
  unsigned wmax;
  int wl, qw;
};

template <int K, int R>
__device__ __forceinline__ TsPick tn_pick(const float (&a)[R][kTsW], const int (&gi)[R]) {
  unsigned key = 0;
  int qb = 0;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const unsigned kq = (gi[q] >= 0) ? __float_as_uint(fabsf(a[q][K])) + 1u : 0u;
    qb = (kq > key) ? q : qb;
    key = (kq > key) ? kq : key;
  }
  TsPick p;
  p.wmax = wave_max_u32(key);
  p.wl = __ffsll((long long)__ballot(key == p.wmax)) - 1;
  p.qw = __builtin_amdgcn_readlane(qb, p.wl);
  return p;
}

// One elimination step (column K) of a node of 64 V R rows, R per lane of V
// waves.  Each wave's candidate for the step (found at the end of the
// previous step's update of column K) is stored by its lane in the LDS slot
// of its wave for this step's parity: key, row number and row (under a
// uniform branch on which of the lane's rows won).  After one barrier every
// lane picks the best of the V slots (ties to the lowest wave) and reads that
// row back by broadcast.  Slots alternate between two parities, so a wave
// writing step K+1's candidate never meets a wave still reading step K's: to
// pass step K+1's barrier every wave must have finished step K.  The update
// does column K+1 first and finds the next candidate from it, so the wave
// maximum's latency overlaps the remaining columns' FMAs.  The step has no
// other branch, so the row registers need no copies between steps: a node
// out of rows elects no row (its winner index reads -1) and changes nothing,
// and columns K >= w are zero, so their steps change nothing either (their
// winners are not used).
template <int K, int R, int V>
__device__ __forceinline__ void tn_step(float (&a)[R][kTsW], int (&gi)[R], int* __restrict__ swin,
                                        TsNodeSmem<V>& sm, float* __restrict__ lu_out, TsPick& pk) {
  constexpr int P = K & 1;
  const int lane = threadIdx.x & 63;
  const int wv = (V > 1) ? __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) : 0;
#pragma unroll
  for (int q = 0; q < R; ++q)
    if (q == pk.qw && lane == pk.wl) {
#pragma unroll
      for (int j = 0; j < kTsW; ++j) sm.row[P][wv][j] = a[q][j];
      sm.idx[P][wv] = gi[q];
      sm.key[P][wv] = pk.wmax;
    }
  int ws = 0;  // the winning wave
  if constexpr (V == 1) {
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the row is in LDS
    __builtin_amdgcn_wave_barrier();
  } else {
    __syncthreads();
    unsigned best = sm.key[P][0];
#pragma unroll
    for (int v = 1; v < V; ++v) {
      const unsigned kv = sm.key[P][v];
      ws = (kv > best) ? v : ws;
      best = (kv > best) ? kv : best;
    }
    ws = __builtin_amdgcn_readfirstlane(ws);
  }
  if (wv == ws) {
#pragma unroll
    for (int q = 0; q < R; ++q) gi[q] = (q == pk.qw && lane == pk.wl) ? -1 : gi[q];
  }
  const float4* s4 = reinterpret_cast<const float4*>(sm.row[P][ws]);
  float r[kTsW];
#pragma unroll
  for (int j4 = K / 4; j4 < kTsW / 4; ++j4) {
    const float4 v = s4[j4];
    r[4 * j4] = v.x;
    r[4 * j4 + 1] = v.y;
    r[4 * j4 + 2] = v.z;
    r[4 * j4 + 3] = v.w;
  }
  const int ib = sm.idx[P][ws];
  if (threadIdx.x == 0) swin[K] = ib;
  // the winner of step K, eliminated by the K earlier winners with its
  // multipliers kept in place, is row K of the winners' LU
  if (lu_out && threadIdx.x < kTsW) lu_out[K * kTsW + threadIdx.x] = sm.row[P][ws][threadIdx.x];
  const float piv = r[K];
  const float rp = (piv != 0.0f) ? __builtin_amdgcn_rcpf(piv) : 0.0f;
  float l[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    l[q] = (gi[q] >= 0) ? a[q][K] * rp : 0.0f;
    a[q][K] = l[q];
  }
  if constexpr (K + 1 < kTsW) {
#pragma unroll
    for (int q = 0; q < R; ++q) a[q][K + 1] -= l[q] * r[K + 1];
    pk = tn_pick<K + 1, R>(a, gi);
  }
#pragma unroll
  for (int q = 0; q < R; ++q)
#pragma unroll
    for (int j = K + 2; j < kTsW; ++j) a[q][j] -= l[q] * r[j];
  if constexpr (V == 1) {
    // one wave: the row has been read before the next step's writes
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
  }
  __builtin_amdgcn_sched_barrier(0);
}

template <int R, int V, int... Ks>
__device__ __forceinline__ void tn_all(std::integer_sequence<int, Ks...>, float (&a)[R][kTsW], int (&gi)[R],
                                       int* __restrict__ swin, TsNodeSmem<V>& sm, float* __restrict__ lu_out) {
  TsPick pk = tn_pick<0, R>(a, gi);
  (tn_step<Ks, R, V>(a, gi, swin, sm, lu_out, pk), ...);
}

// Loads the 32-column segment [c0, c0+32) of row g (zero beyond column w).
// c0 is a multiple of 32 and rows are ld >= roundup(N, 64) floats long, so the
// segment is 128-byte aligned and inside the row.
__device__ __forceinline__ void ts_load_row(const float* __restrict__ W, long long ld, long long g, long long c0,
                                            int w, float (&r)[kTsW]) {
  const float4* src = reinterpret_cast<const float4*>(W + g * ld + c0);
#pragma unroll
  for (int j4 = 0; j4 < kTsW / 4; ++j4) {
    const float4 v = src[j4];
    r[4 * j4 + 0] = (4 * j4 + 0 < w) ? v.x : 0.0f;
    r[4 * j4 + 1] = (4 * j4 + 1 < w) ? v.y : 0.0f;
    r[4 * j4 + 2] = (4 * j4 + 2 < w) ? v.z : 0.0f;
    r[4 * j4 + 3] = (4 * j4 + 3 < w) ? v.w : 0.0f;
  }
}

// cand: the candidate lists of every level (ts_tree_sizes entries); cnt: the
// group counters, zero between leaves (the last arriver resets its own).
// Outputs of the root: lu_out (w x w winners' LU, stride 32), top_out (the
// original top rows' leaf columns, 32 x 32), meta, ipiv[c0 .. c0+w).
// 64 V threads per workgroup; a node's row slot t (0 .. 64 V R) is held by
// thread t % (64 V) as its row t / (64 V).
template <int R, int V>
__global__ __launch_bounds__(64 * V) void k_tslu_tree(const float* __restrict__ W, long long ld, long long N,
                                                      long long c0, int w, int* cand, unsigned* cnt,
                                                      float* __restrict__ lu_out, float* __restrict__ top_out,
                                                      int* __restrict__ meta, int* __restrict__ ipiv) {
  using Nd = TsNode<R, V>;
  constexpr int T = Nd::threads;
  __shared__ int swin[kTsW];
  __shared__ __attribute__((aligned(16))) TsNodeSmem<V> sm;
  __shared__ int slast;
  __builtin_amdgcn_s_setprio(3);
  const int tid = threadIdx.x;
  const long long m = N - c0;
  long long nl = (m + Nd::rows - 1) / Nd::rows;  // nodes of the current level
  long long node = blockIdx.x;
  long long off = 0;   // first candidate entry of the current level
  long long coff = 0;  // first counter of the current level's groups
  int gi[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const long long g = c0 + node * Nd::rows + tid + T * q;
    gi[q] = (g < N) ? (int)g : -1;
  }
  while (true) {
    float a[R][kTsW];
#pragma unroll
    for (int q = 0; q < R; ++q) ts_load_row(W, ld, gi[q] >= 0 ? gi[q] : c0, c0, w, a[q]);
#pragma unroll
    for (int q = 0; q < R; ++q)
#pragma unroll
      for (int j = 0; j < kTsW; ++j) a[q][j] = (gi[q] >= 0) ? a[q][j] : 0.0f;
    const bool root = (nl == 1);
    tn_all<R, V>(std::make_integer_sequence<int, kTsW>{}, a, gi, swin, sm, root ? lu_out : nullptr);
    __syncthreads();
    if (root) break;
    if (tid < kTsW) cand[off + node * kTsW + tid] = (tid < w) ? swin[tid] : -1;
    // hand the winners on: the group's last node to finish runs the parent
    const long long grp = node / Nd::fan;
    const unsigned gsz = (unsigned)min((long long)Nd::fan, nl - Nd::fan * grp);
    __syncthreads();
    if (tid == 0) {
      __threadfence();
      const unsigned old = atomicAdd(&cnt[coff + grp], 1u);
      const int last = (old == gsz - 1) ? 1 : 0;
      if (last) atomicExch(&cnt[coff + grp], 0u);
      slast = last;
    }
    __syncthreads();
    if (!slast) return;
    __threadfence();
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int t = tid + T * q;
      int ng = -1;
      if (t < (int)(kTsW * gsz)) {
        ng = __hip_atomic_load(&cand[off + Nd::fan * grp * kTsW + t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (ng < c0 || ng >= N) ng = -1;
      }
      gi[q] = ng;
    }
    off += nl * kTsW;
    coff += (nl + Nd::fan - 1) / Nd::fan;
    nl = (nl + Nd::fan - 1) / Nd::fan;
    node = grp;
  }
  // root: the original top rows' leaf columns, then the interchange sequence
  // that brings the winners p_0..p_{w-1} to rows c0..c0+w-1 in order
  // (ipiv[c0+k] = current position of p_k), formed over the 64 positions it
  // can touch: lanes 0..31 the top rows, lane 32+k the original position of
  // p_k when that lies below the top rows.
  if (tid >= 64) return;
#pragma unroll
  for (int i = tid; i < kTsW * kTsW / 4; i += 64) {
    const int r = i >> 3, j4 = (i & 7) * 4;
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (r < w) v = *reinterpret_cast<const float4*>(W + (c0 + r) * ld + c0 + j4);
    *reinterpret_cast<float4*>(top_out + r * kTsW + j4) = v;
  }
  {
    const int lane = tid;
    int pos = -1;
    if (lane < w) {
      pos = (int)c0 + lane;
    } else if (lane >= 32 && lane - 32 < w) {
      const int p = swin[lane - 32];
      if (p >= c0 + w && p < N) pos = p;
    }
    int content = pos;
    for (int k = 0; k < w; ++k) {
      const int t = (int)c0 + k, p = swin[k];
      const unsigned long long bp = __ballot(p >= 0 && content == p);
      int cur = t;
      if (bp) {
        const int lp = __ffsll((long long)bp) - 1;
        cur = __builtin_amdgcn_readlane(pos, lp);
        if (cur != t) {
          const int ct = __builtin_amdgcn_readlane(content, k);
          if (lane == lp) content = ct;
          if (lane == k) content = p;
        }
      }
      if (lane == 0) {
        meta[k] = cur;
        ipiv[t] = cur;
      }
    }
    const unsigned long long bd = __ballot(lane >= 32 && pos >= 0 && content != pos);
    const int rank = __popcll(bd & ((1ull << lane) - 1ull));
    if (lane >= 32 && pos >= 0 && content != pos) {
      meta[32 + rank] = pos;
      meta[64 + rank] = content;
    }
    if (lane == 0) meta[96] = __popcll(bd);
    if (lane < kTsW) meta[128 + lane] = (lane < w) ? content : -1;
  }
}

// Running maximum of |v| that keeps NaN: for non-negative floats the unsigned
// order of the bit patterns is the numeric order with NaN above +inf
// (fmaxf would drop a NaN operand).
__device__ __forceinline__ float abs_max_keep_nan(float g, float v) {
  return __uint_as_float(max(__float_as_uint(g), __float_as_uint(fabsf(v))));
}

// Largest magnitude of a float as its bit pattern (NaN above +inf), folded
// into *growth; called by whole waves.
__device__ __forceinline__ void ts_growth_max(float v, unsigned* growth) {
  unsigned b = __float_as_uint(fabsf(v));
  b = wave_max_u32(b);
  if ((threadIdx.x & 63) == 0) atomicMax(growth, b);
}

// The same for a whole 256-thread workgroup with one atomic: every thread of
// the workgroup must call it (it holds a barrier).  Atomics of every wave on
// the one word serialise at its L2 channel (a 70,000-row finish spent ~8 us
// on them).
__device__ __forceinline__ void ts_growth_max_wg(float v, unsigned* growth, unsigned (&wmax)[4]) {
  const unsigned b = wave_max_u32(__float_as_uint(fabsf(v)));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = b;
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(growth, max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3])));
}

// L21 = A21 U11^-1 (one row per thread, U11 read from LDS by broadcast),
// against the root's winners' LU, which workgroup 0 also writes into the top
// rows; then, in workgroups row_blocks.., the permutation on the panel's
// columns [p0, p1) outside the leaf; the last workgroup inverts the leaf's
// unit lower block into linv_out.  No workgroup reads what another writes:
// rows replaced by the interchanges read the root's copy of the top rows.
template <int K>
__device__ __forceinline__ void tf_solve_step(float (&x)[kTsW], int w, const float (*__restrict__ sU)[kTsW],
                                              const float* __restrict__ sR) {
  if (K < w) {
    const float4* u4 = reinterpret_cast<const float4*>(sU[K]);
    float u[kTsW];
#pragma unroll
    for (int j4 = K / 4; j4 < kTsW / 4; ++j4) {
      const float4 v = u4[j4];
      u[4 * j4] = v.x;
      u[4 * j4 + 1] = v.y;
      u[4 * j4 + 2] = v.z;
      u[4 * j4 + 3] = v.w;
    }
    x[K] *= sR[K];
#pragma unroll
    for (int j = K + 1; j < kTsW; ++j) x[j] -= x[K] * u[j];
  }
}

template <int... Ks>
__device__ __forceinline__ void tf_solve(std::integer_sequence<int, Ks...>, float (&x)[kTsW], int w,
                                         const float (*__restrict__ sU)[kTsW], const float* __restrict__ sR) {
  (tf_solve_step<Ks>(x, w, sU, sR), ...);
}

__global__ __launch_bounds__(256) void k_tslu_finish(float* __restrict__ W, long long ld, long long N, long long c0,
                                                     int w, long long p0, long long p1,
                                                     const float* __restrict__ lu_in,
                                                     const float* __restrict__ top_in, const int* __restrict__ meta,
                                                     int row_blocks, float* __restrict__ linv_out) {
  __shared__ __attribute__((aligned(16))) float sU[kTsW][kTsW];
  __builtin_amdgcn_s_setprio(3);
  __shared__ float sR[kTsW];
  __shared__ int sdpos[kTsW], sdsrc[kTsW], stop[kTsW];
  __shared__ int snd;
  const int tid = threadIdx.x;
  if (tid < kTsW) {
    const int dp = meta[32 + tid], ds = meta[64 + tid], ts = meta[128 + tid];
    sdpos[tid] = dp;
    sdsrc[tid] = (ds >= c0 && ds < c0 + w && dp >= c0 + w && dp < N) ? ds : -1;
    stop[tid] = (tid < w && ts >= c0 && ts < N) ? ts : (int)c0 + tid;
  }
  if (tid == 0) snd = min(max(meta[96], 0), kTsW);
  if (blockIdx.x == gridDim.x - 1) {
    // the last workgroup: inverse of the winners' unit lower factor
    // (identity beyond w), one column per thread by forward substitution,
    // for the later triangular solves with this block
    __shared__ float sL[kTsW][kTsW + 1];
    for (int i = tid; i < kTsW * kTsW; i += 256) {
      const int r = i / kTsW, j = i % kTsW;
      sL[r][j] = (j < r && r < w) ? lu_in[i] : 0.0f;
    }
    __syncthreads();
    if (tid < kTsW) {
      // column tid in registers (entries above the diagonal come out zero);
      // the LDS reads of a row do not depend on the column and go together
      float y[kTsW];
#pragma unroll
      for (int r = 0; r < kTsW; ++r) {
        float acc = (r == tid) ? 1.0f : 0.0f;
#pragma unroll
        for (int k = 0; k < r; ++k) acc -= sL[r][k] * y[k];
        y[r] = acc;
      }
#pragma unroll
      for (int r = 0; r < kTsW; ++r) linv_out[r * kTsW + tid] = y[r];
    }
    return;
  }
  if ((int)blockIdx.x >= row_blocks) {
    __syncthreads();
    const long long c = p0 + (long long)(blockIdx.x - row_blocks) * 256 + tid;
    if (c >= p1 || (c >= c0 && c < c0 + w)) return;
    const int nd = snd;
    float vt[kTsW], vd[kTsW];
#pragma unroll
    for (int k = 0; k < kTsW; ++k) {
      vt[k] = (k < w) ? W[(long long)stop[k] * ld + c] : 0.0f;
      vd[k] = (k < nd && sdsrc[k] >= 0) ? W[(long long)sdsrc[k] * ld + c] : 0.0f;
    }
    // every read of the column lands before its first write
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
#pragma unroll
    for (int k = 0; k < kTsW; ++k) {
      if (k < w) W[(c0 + k) * ld + c] = vt[k];
      if (k < nd && sdsrc[k] >= 0) W[(long long)sdpos[k] * ld + c] = vd[k];
    }
    return;
  }
  {
    float v[kTsW * kTsW / 256];
#pragma unroll
    for (int q = 0; q < kTsW * kTsW / 256; ++q) {
      const int i = tid + 256 * q, r = i / kTsW, j = i % kTsW;
      v[q] = 0.0f;
      if (r < w && j < w) v[q] = lu_in[i];
    }
#pragma unroll
    for (int q = 0; q < kTsW * kTsW / 256; ++q) {
      const int i = tid + 256 * q, r = i / kTsW, j = i % kTsW;
      sU[r][j] = (j >= r) ? v[q] : 0.0f;
      if (blockIdx.x == 0 && r < w && j < w) W[(c0 + r) * ld + c0 + j] = v[q];
    }
  }
  __syncthreads();
  if (tid < kTsW) {
    const float d = sU[tid][tid];
    sR[tid] = (tid < w && d != 0.0f) ? 1.0f / d : 0.0f;
  }
  __syncthreads();
  const long long r = c0 + w + (long long)blockIdx.x * 256 + tid;
  if (r >= N) return;
  int src = -1;
  const int nd = snd;
  for (int i = 0; i < nd; ++i)
    if (sdpos[i] == r) src = sdsrc[i];
  float x[kTsW];
  if (src >= 0) {
    const float4* t4 = reinterpret_cast<const float4*>(top_in + (src - c0) * kTsW);
#pragma unroll
    for (int j4 = 0; j4 < kTsW / 4; ++j4) {
      const float4 t = t4[j4];
      x[4 * j4] = t.x;
      x[4 * j4 + 1] = t.y;
      x[4 * j4 + 2] = t.z;
      x[4 * j4 + 3] = t.w;
    }
  } else {
    ts_load_row(W, ld, r, c0, w, x);
  }
  tf_solve(std::make_integer_sequence<int, kTsW>{}, x, w, sU, sR);
  float* dst = W + r * ld + c0;
  if (w == kTsW) {
    // the segment is 128-byte aligned (c0 a multiple of 32, ld of 64)
#pragma unroll
    for (int j4 = 0; j4 < kTsW / 4; ++j4)
      *reinterpret_cast<float4*>(dst + 4 * j4) = make_float4(x[4 * j4], x[4 * j4 + 1], x[4 * j4 + 2], x[4 * j4 + 3]);
  } else {
#pragma unroll
    for (int j = 0; j < kTsW; ++j)
      if (j < w) dst[j] = x[j];
  }
}

// ---------------------------------------------------------------------------
// Leaves with diagonal pivots (the factorisations that try them first), 64
// columns wide: a wave holds the leaf's 64 top rows, one per lane, and the
// recursion above has one level fewer than with 32-column leaves.
constexpr int kDW = 64;

// Loads the 64-column segment [c0, c0+64) of row g (zero beyond column w).
__device__ __forceinline__ void dw_load_row(const float* __restrict__ W, long long ld, long long g, long long c0,
                                            int w, float (&r)[kDW]) {
  const float4* src = reinterpret_cast<const float4*>(W + g * ld + c0);
#pragma unroll
  for (int j4 = 0; j4 < kDW / 4; ++j4) {
    const float4 v = src[j4];
    r[4 * j4 + 0] = (4 * j4 + 0 < w) ? v.x : 0.0f;
    r[4 * j4 + 1] = (4 * j4 + 1 < w) ? v.y : 0.0f;
    r[4 * j4 + 2] = (4 * j4 + 2 < w) ? v.z : 0.0f;
    r[4 * j4 + 3] = (4 * j4 + 3 < w) ? v.w : 0.0f;
  }
}

// One step (column K) of the top block's elimination: lane K's row is final
// and reaches the other lanes by readlane broadcasts (no LDS round trip per
// step).  Rows go to memory after the last step, all lanes at once.
template <int K>
__device__ __forceinline__ void dw_step(float (&a)[kDW], int w, float& gmax) {
  if (K < w) {
    const int lane = threadIdx.x;
    const float piv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a[K]), K));
    if (piv == 0.0f) gmax = abs_max_keep_nan(gmax, __int_as_float(0x7f800000));
    const float rp = (piv != 0.0f) ? __builtin_amdgcn_rcpf(piv) : 0.0f;
    const bool act = lane > K && lane < w;
    const float l = a[K] * rp;
    if (act) {
      a[K] = l;
      gmax = abs_max_keep_nan(gmax, l);
    }
#pragma unroll
    for (int j = K + 1; j < kDW; ++j) {
      const float u = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a[j]), K));
      if (act) a[j] -= l * u;
    }
  }
  __builtin_amdgcn_sched_barrier(0);
}

template <int... Ks>
__device__ __forceinline__ void dw_all(std::integer_sequence<int, Ks...>, float (&a)[kDW], int w, float& gmax) {
  (dw_step<Ks>(a, w, gmax), ...);
}

// The top w x w block (w <= 64) of a diagonal-pivot leaf, one wave: its LU
// (multipliers below the diagonal) to lu_out (64 x 64), the inverses of the
// unit lower factor's two 32 x 32 diagonal blocks to linv_out (two 32 x 32
// blocks, for the strip solves), ipiv = identity, and the largest multiplier
// |l| (a zero pivot counts as +inf) folded into *growth.
__global__ __launch_bounds__(64) void k_leaf_diag(const float* __restrict__ W, long long ld, long long N,
                                                  long long c0, int w, float* __restrict__ lu_out,
                                                  float* __restrict__ linv_out, int* __restrict__ ipiv,
                                                  unsigned* __restrict__ growth) {
  __shared__ float sL[kDW][kDW + 1];
  __builtin_amdgcn_s_setprio(3);
  const int lane = threadIdx.x;
  float a[kDW];
  dw_load_row(W, ld, (lane < w) ? c0 + lane : c0, c0, w, a);
  if (lane >= w)
#pragma unroll
    for (int j = 0; j < kDW; ++j) a[j] = 0.0f;
  float gmax = 0.0f;
  dw_all(std::make_integer_sequence<int, kDW>{}, a, w, gmax);
  // lane r's registers hold row r of the LU (multipliers left of the diagonal)
  if (lane < w)
#pragma unroll
    for (int j4 = 0; j4 < kDW / 4; ++j4)
      *reinterpret_cast<float4*>(lu_out + lane * kDW + 4 * j4) =
          make_float4(a[4 * j4], a[4 * j4 + 1], a[4 * j4 + 2], a[4 * j4 + 3]);
  ts_growth_max(gmax, growth);
  if (lane < w) ipiv[c0 + lane] = (int)c0 + lane;
  // the unit lower factor (identity beyond w) for the inverses: lane r's
  // multipliers are its row of L
#pragma unroll
  for (int j = 0; j < kDW; ++j) sL[lane][j] = (j < lane && lane < w) ? a[j] : 0.0f;
  __builtin_amdgcn_s_waitcnt(0xc07f);
  __builtin_amdgcn_wave_barrier();
  // lane t: column t % 32 of the inverse of diagonal block t / 32
  const int blk = lane >> 5, cc = lane & 31, o = 32 * blk;
  float y[32];
#pragma unroll
  for (int r = 0; r < 32; ++r) {
    float acc = (r == cc) ? 1.0f : 0.0f;
#pragma unroll
    for (int k = 0; k < r; ++k) acc -= sL[o + r][o + k] * y[k];
    y[r] = acc;
  }
#pragma unroll
  for (int r = 0; r < 32; ++r) linv_out[blk * 1024 + r * 32 + cc] = y[r];
}

template <int K>
__device__ __forceinline__ void dwf_step(float (&x)[kDW], int w, const float (*__restrict__ sU)[kDW],
                                         const float* __restrict__ sR) {
  if (K < w) {
    x[K] *= sR[K];
    const float4* u4 = reinterpret_cast<const float4*>(sU[K]);
#pragma unroll
    for (int j4 = (K + 1) / 4; j4 < kDW / 4; ++j4) {
      const float4 v = u4[j4];
      if (4 * j4 > K) x[4 * j4] -= x[K] * v.x;
      if (4 * j4 + 1 > K) x[4 * j4 + 1] -= x[K] * v.y;
      if (4 * j4 + 2 > K) x[4 * j4 + 2] -= x[K] * v.z;
      if (4 * j4 + 3 > K) x[4 * j4 + 3] -= x[K] * v.w;
    }
  }
}

template <int... Ks>
__device__ __forceinline__ void dwf_all(std::integer_sequence<int, Ks...>, float (&x)[kDW], int w,
                                        const float (*__restrict__ sU)[kDW], const float* __restrict__ sR) {
  (dwf_step<Ks>(x, w, sU, sR), ...);
}

// L21 = A21 U11^-1 for the rows below a diagonal-pivot leaf's top block, one
// thread per row, U11 from lu_in by LDS broadcast; workgroup 0 also writes the
// block's LU into W's top rows.  The largest |l| goes to *growth.
__global__ __launch_bounds__(256) void k_finish_diag(float* __restrict__ W, long long ld, long long N, long long c0,
                                                     int w, const float* __restrict__ lu_in,
                                                     unsigned* __restrict__ growth) {
  __shared__ __attribute__((aligned(16))) float sU[kDW][kDW];
  __shared__ float sR[kDW];
  __builtin_amdgcn_s_setprio(3);
  const int tid = threadIdx.x;
  // this thread's row first, then the block's LU (all loads in flight before
  // the LDS stores and barriers)
  const long long r = c0 + w + (long long)blockIdx.x * 256 + tid;
  float x[kDW];
  if (r < N) dw_load_row(W, ld, r, c0, w, x);
  float lv[kDW * kDW / 256];
#pragma unroll
  for (int q = 0; q < kDW * kDW / 256; ++q) {
    const int i = tid + 256 * q, ri = i / kDW, j = i % kDW;
    lv[q] = (ri < w && j < w) ? lu_in[i] : 0.0f;
  }
#pragma unroll
  for (int q = 0; q < kDW * kDW / 256; ++q) {
    const int i = tid + 256 * q, ri = i / kDW, j = i % kDW;
    sU[ri][j] = (j >= ri) ? lv[q] : 0.0f;
    if (blockIdx.x == 0 && ri < w && j < w) W[(c0 + ri) * ld + c0 + j] = lv[q];
  }
  __syncthreads();
  if (tid < kDW) {
    const float d = sU[tid][tid];
    sR[tid] = (tid < w && d != 0.0f) ? 1.0f / d : 0.0f;
  }
  __syncthreads();
  float g = 0.0f;
  if (r < N) {
    dwf_all(std::make_integer_sequence<int, kDW>{}, x, w, sU, sR);
#pragma unroll
    for (int j = 0; j < kDW; ++j) g = abs_max_keep_nan(g, x[j]);
    float* dst = W + r * ld + c0;
    if (w == kDW) {
#pragma unroll
      for (int j4 = 0; j4 < kDW / 4; ++j4)
        *reinterpret_cast<float4*>(dst + 4 * j4) = make_float4(x[4 * j4], x[4 * j4 + 1], x[4 * j4 + 2], x[4 * j4 + 3]);
    } else {
#pragma unroll
      for (int j = 0; j < kDW; ++j)
        if (j < w) dst[j] = x[j];
    }
  }
  __shared__ unsigned wmax[4];
  ts_growth_max_wg(g, growth, wmax);
}

// The kernels above for a leaf of exactly kDW columns (every leaf but the
// last of the matrix), where no column or row test is left at run time.
// Without the per-column width masks the 63 pivot-row broadcasts of a step
// go to distinct scalar registers and pipeline (the generic leaf funnels them
// through one, 25.5 -> 15.4 us per leaf alone), and the finish reads U11's
// rows as wave-uniform values (scalar loads) instead of through LDS.  Results
// are bit-identical to the generic kernels'.
template <int K>
__device__ __forceinline__ void dw_step64(float (&a)[kDW], float& gmax) {
  const int lane = threadIdx.x;
  const float piv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a[K]), K));
  if (piv == 0.0f) gmax = abs_max_keep_nan(gmax, __int_as_float(0x7f800000));
  const float rp = (piv != 0.0f) ? __builtin_amdgcn_rcpf(piv) : 0.0f;
  const bool act = lane > K;
  const float l = a[K] * rp;
  if (act) {
    a[K] = l;
    gmax = abs_max_keep_nan(gmax, l);
  }
  float u[kDW];
#pragma unroll
  for (int j = K + 1; j < kDW; ++j) u[j] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a[j]), K));
#pragma unroll
  for (int j = K + 1; j < kDW; ++j)
    if (act) a[j] -= l * u[j];
  __builtin_amdgcn_sched_barrier(0);
}

template <int... Ks>
__device__ __forceinline__ void dw_all64(std::integer_sequence<int, Ks...>, float (&a)[kDW], float& gmax) {
  (dw_step64<Ks>(a, gmax), ...);
}

__global__ __launch_bounds__(64) void k_leaf_diag64(const float* __restrict__ W, long long ld, long long c0,
                                                    float* __restrict__ lu_out, float* __restrict__ linv_out,
                                                    int* __restrict__ ipiv, unsigned* __restrict__ growth) {
  __shared__ float sL[kDW][kDW + 1];
  __builtin_amdgcn_s_setprio(3);
  const int lane = threadIdx.x;
  float a[kDW];
  dw_load_row(W, ld, c0 + lane, c0, kDW, a);
  float gmax = 0.0f;
  dw_all64(std::make_integer_sequence<int, kDW>{}, a, gmax);
#pragma unroll
  for (int j4 = 0; j4 < kDW / 4; ++j4)
    *reinterpret_cast<float4*>(lu_out + lane * kDW + 4 * j4) =
        make_float4(a[4 * j4], a[4 * j4 + 1], a[4 * j4 + 2], a[4 * j4 + 3]);
  ts_growth_max(gmax, growth);
  ipiv[c0 + lane] = (int)c0 + lane;
#pragma unroll
  for (int j = 0; j < kDW; ++j) sL[lane][j] = (j < lane) ? a[j] : 0.0f;
  __builtin_amdgcn_s_waitcnt(0xc07f);
  __builtin_amdgcn_wave_barrier();
  const int blk = lane >> 5, cc = lane & 31, o = 32 * blk;
  float y[32];
#pragma unroll
  for (int r = 0; r < 32; ++r) {
    float acc = (r == cc) ? 1.0f : 0.0f;
#pragma unroll
    for (int k = 0; k < r; ++k) acc -= sL[o + r][o + k] * y[k];
    y[r] = acc;
  }
#pragma unroll
  for (int r = 0; r < 32; ++r) linv_out[blk * 1024 + r * 32 + cc] = y[r];
}

template <int K>
__device__ __forceinline__ void dws_step64(float (&x)[kDW], const float* __restrict__ lu) {
  const float d = lu[K * kDW + K];
  x[K] *= (d != 0.0f) ? 1.0f / d : 0.0f;
#pragma unroll
  for (int j = K + 1; j < kDW; ++j) x[j] -= x[K] * lu[K * kDW + j];
}

template <int... Ks>
__device__ __forceinline__ void dws_all64(std::integer_sequence<int, Ks...>, float (&x)[kDW],
                                          const float* __restrict__ lu) {
  (dws_step64<Ks>(x, lu), ...);
}

__global__ __launch_bounds__(256) void k_finish_diag64(float* __restrict__ W, long long ld, long long N, long long c0,
                                                       const float* __restrict__ lu_in, unsigned* __restrict__ growth) {
  __shared__ unsigned wmax[4];
  __builtin_amdgcn_s_setprio(3);
  const int tid = threadIdx.x;
  const long long r = c0 + kDW + (long long)blockIdx.x * 256 + tid;
  float x[kDW];
  if (r < N) dw_load_row(W, ld, r, c0, kDW, x);
  if (blockIdx.x == 0)
#pragma unroll
    for (int q = 0; q < kDW * kDW / 256; ++q) {
      const int i = tid + 256 * q;
      W[(c0 + i / kDW) * ld + c0 + i % kDW] = lu_in[i];
    }
  float g = 0.0f;
  if (r < N) {
    dws_all64(std::make_integer_sequence<int, kDW>{}, x, lu_in);
#pragma unroll
    for (int j = 0; j < kDW; ++j) g = abs_max_keep_nan(g, x[j]);
    float* dst = W + r * ld + c0;
#pragma unroll
    for (int j4 = 0; j4 < kDW / 4; ++j4)
      *reinterpret_cast<float4*>(dst + 4 * j4) = make_float4(x[4 * j4], x[4 * j4 + 1], x[4 * j4 + 2], x[4 * j4 + 3]);
  }
  ts_growth_max_wg(g, growth, wmax);
}

}  // namespace qd
