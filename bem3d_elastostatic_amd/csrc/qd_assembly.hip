// qd_assembly.hip -- element Gauss geometry and the ghmat8 collocation
// assembly on gfx950.
//
// Work decomposition (assembly_mod.f90:53-223 restated for the GPU):
//  * regular pairs (collocation node not in the element): one lane per
//    collocation node, 64 consecutive nodes per wave; a wave walks a chunk of
//    consecutive elements.  Element data (Gauss table, codes, values) are
//    wave-uniform and come through scalar loads.  The H integrals
//    (Kelvin traction kernel, 9 components) and the G integrals (displacement
//    kernel, 6 components) run as separate passes: this halves the
//    accumulator registers, and the G pass only visits elements whose G
//    terms land somewhere (a displacement-prescribed dof or a non-zero
//    prescribed traction).
//  * chunks are coloured so that no two chunks of one launch share a node;
//    the column-block updates of A (<= 8 contributions per block) are then
//    plain read-modify-write stores with no atomics and a fixed order.
//  * the rigid-body diagonal (d1) and the FI terms are summed per lane over
//    the chunk and written as per-chunk partials; a per-node finalize kernel
//    reduces them in chunk order, adds the singular-element terms (locin8,
//    computed by a separate kernel) and writes the diagonal block and FI.
#include <hip/hip_runtime.h>

#include "qd_internal.h"
#include "qd_device.h"

namespace qd {

// assembly_mod.f90:6-19 (literals verbatim)
__constant__ double c_gi[6] = {0.238619186083197, -0.238619186083197, 0.661209386466265,
                               -0.661209386466265, 0.932469514203152, -0.932469514203152};
__constant__ double c_ome[6] = {0.467913934572691, 0.467913934572691, 0.360761573048139,
                                0.360761573048139, 0.171324492379170, 0.171324492379170};
__constant__ double c_rr[4] = {0.0694318422, 0.3300094782, 0.6699905218, 0.9305681558};
__constant__ double c_ss[4] = {0.0571041961, 0.2768430136, 0.5835904324, 0.8602401357};
__constant__ double c_omei[4] = {0.1739274226, 0.3260725774, 0.3260725774, 0.1739274226};
__constant__ double c_omej[4] = {0.1355069134, 0.2034645680, 0.1298475476, 0.0311809709};

// precompute_element_geometry / etaxja (assembly_mod.f90:28-51, 225-337):
// one thread per (element, Gauss slot).  asmtab32 holds the same point's
// normal and weight products in FP32 for the far-field passes.
__global__ void k_geometry(int ne, const int* __restrict__ nconec, const int* __restrict__ kconec0,
                           const double* __restrict__ x, const double* __restrict__ y,
                           const double* __restrict__ z, double* __restrict__ geo,
                           double* __restrict__ asmtab, float* __restrict__ asmtab32) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= (long long)ne * kGP) return;
  const int e = (int)(tid / kGP), k = (int)(tid % kGP);
  const int nc = nconec[e];
  const int npg = nc < 9 ? 4 : 6;
  const size_t blk = (size_t)ne * kGP;
  double out[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // eta1 eta2 eta3 xja xg1 xg2 xg3
  double f[9], fd1[9], fd2[9];
  double p12[9];
  for (int m = 0; m < 9; ++m) f[m] = p12[m] = 0.0;
  if (k < npg * npg) {
    const int i = k / npg, j = k % npg;
    double xj1[9], xj2[9], xj3[9];
    for (int m = 0; m < 9; ++m) {
      int nd = m < nc ? kconec0[e * 9 + m] : 0;
      xj1[m] = m < nc ? x[nd] : 0.0;
      xj2[m] = m < nc ? y[nd] : 0.0;
      xj3[m] = m < nc ? z[nd] : 0.0;
    }
    double g1, g2;
    if (nc == 9) {
      g1 = c_gi[i];
      g2 = c_gi[j];
    } else {
      g1 = c_ss[j];
      g2 = c_rr[i] * (1.0 - c_ss[j]);
    }
    shape(nc, g1, g2, f, fd1, fd2);
    double xg1 = 0.0, xg2 = 0.0, xg3 = 0.0;
    for (int m = 0; m < nc; ++m) {
      xg1 += f[m] * xj1[m];
      xg2 += f[m] * xj2[m];
      xg3 += f[m] * xj3[m];
    }
    double s1, s2, s3;
    normal_dc(nc, fd1, fd2, xj1, xj2, xj3, s1, s2, s3);
    double xja = sqrt(s1 * s1 + s2 * s2 + s3 * s3);
    out[0] = s1 / xja;
    out[1] = s2 / xja;
    out[2] = s3 / xja;
    out[3] = xja;
    out[4] = xg1;
    out[5] = xg2;
    out[6] = xg3;
    // extin8's weight product, in its evaluation order (assembly_mod.f90:363-369)
    for (int m = 0; m < nc; ++m)
      p12[m] = (nc == 9) ? f[m] * xja * c_ome[i] * c_ome[j] : f[m] * xja * c_omei[i] * c_omej[j];
  }
  const size_t slot = (size_t)e * kGP + k;
  for (int b = 0; b < 7; ++b) geo[b * blk + slot] = out[b];
  for (int m = 0; m < 9; ++m) geo[7 * blk + slot * 9 + m] = f[m];
  double* t = asmtab + slot * kAsmStride;
  t[0] = out[4];
  t[1] = out[5];
  t[2] = out[6];
  t[3] = out[0];
  t[4] = out[1];
  t[5] = out[2];
  for (int m = 0; m < 9; ++m) t[6 + m] = p12[m];
  t[15] = 0.0;
  float* t32 = asmtab32 + slot * kAsmStride32;
  for (int c = 0; c < 3; ++c) t32[c] = (float)out[c];
  for (int m = 0; m < 9; ++m) t32[3 + m] = (float)p12[m];
  for (int c = 12; c < kAsmStride32; ++c) t32[c] = 0.0f;
}

// Far-field test data per element: centroid (mean of the nodes) and the
// squared distance beyond which a collocation node is more than two element
// diameters (twice the largest node-to-centroid distance) away.
__global__ void k_elem_far(int ne, const int* __restrict__ nconec, const int* __restrict__ kconec0,
                           const double* __restrict__ x, const double* __restrict__ y,
                           const double* __restrict__ z, double* __restrict__ far) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int nc = nconec[e];
  double c0 = 0.0, c1 = 0.0, c2 = 0.0;
  for (int m = 0; m < nc; ++m) {
    const int nd = kconec0[e * 9 + m];
    c0 += x[nd];
    c1 += y[nd];
    c2 += z[nd];
  }
  c0 /= nc;
  c1 /= nc;
  c2 /= nc;
  double r2 = 0.0;
  for (int m = 0; m < nc; ++m) {
    const int nd = kconec0[e * 9 + m];
    const double d0 = x[nd] - c0, d1 = y[nd] - c1, d2 = z[nd] - c2;
    r2 = fmax(r2, d0 * d0 + d1 * d1 + d2 * d2);
  }
  // distance > 2 diameters = 4 radii
  far[4 * e] = c0;
  far[4 * e + 1] = c1;
  far[4 * e + 2] = c2;
  far[4 * e + 3] = 16.0 * r2;
}

// The collocation node (xp, yp, zp) is in element e's far field.
__device__ __forceinline__ bool is_far(const double* __restrict__ far, int e, double xp, double yp, double zp) {
  const double d0 = xp - far[4 * e], d1 = yp - far[4 * e + 1], d2 = zp - far[4 * e + 2];
  return d0 * d0 + d1 * d1 + d2 * d2 > far[4 * e + 3];
}

struct AsmArgs {
  int n, ne;
  long long lda;
  const double* x;
  const double* y;
  const double* z;
  const double* asmtab;
  const ElemInfo* einfo;
  const double* ebc;
  const double* evp;
  const int* chunk_start;
  double* A;
  double* dpart;
  double* fpart;
  double mu, c3, c4, c5, c6;
  int write_A;
  const float* asmtab32;  // far-field passes: FP32 normals and weight products
  const double* elfar;    // far-field passes: k_elem_far's table
  const unsigned* elop;   // write plan, kElemOps words per element
};

// H pass: h_elem of extin8 (assembly_mod.f90:339-384) with funkelvin's
// traction kernel (618-650), then ghmat8's H terms (100-190).  With FAR32,
// a lane whose collocation node is in the element's far field evaluates the
// kernel and sums the element integral in FP32 (r itself is formed in FP64);
// the other lanes, and everything after the integral, stay FP64.
template <int NC, bool FAR32>
__global__ __launch_bounds__(64) void k_hpass(AsmArgs a, const int* __restrict__ chunks, int coff,
                                               const double* __restrict__ asmtab, const ElemInfo* __restrict__ einfo,
                                               const double* __restrict__ ebc_all, const double* __restrict__ evp_all,
                                               double* __restrict__ Aout) {
  constexpr int NG = (NC == 9) ? 36 : 16;
  __shared__ double carry[kCarrySlots][9][64];  // column blocks carried between a chunk's elements
  const int chunk = chunks[coff + blockIdx.y];
  const int e0 = a.chunk_start[chunk], e1 = a.chunk_start[chunk + 1];
  const int in = blockIdx.x * 64 + threadIdx.x;
  const bool valid = in < a.n;
  const int inr = valid ? in : a.n - 1;
  const double xp = a.x[inr], yp = a.y[inr], zp = a.z[inr];
  const double c5 = a.c5, c6 = a.c6;
  double fi0 = 0.0, fi1 = 0.0, fi2 = 0.0;
  double d1[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) d1[i] = 0.0;

  for (int e = e0; e < e1; ++e) {
    const ElemInfo* inf = einfo + e;
    bool sing = false;
#pragma unroll
    for (int k = 0; k < NC; ++k) sing |= (inf->node[k] == in);
    double acc[NC][3][3];  // acc[kk][ii][jj] = h(jj, 3kk+ii)
#pragma unroll
    for (int kk = 0; kk < NC; ++kk)
#pragma unroll
      for (int ii = 0; ii < 3; ++ii)
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) acc[kk][ii][jj] = 0.0;
    const double* gp = asmtab + (size_t)e * kGP * kAsmStride;
    if (FAR32 && is_far(a.elfar, e, xp, yp, zp)) {
      float fa[NC][3][3];
#pragma unroll
      for (int kk = 0; kk < NC; ++kk)
#pragma unroll
        for (int ii = 0; ii < 3; ++ii)
#pragma unroll
          for (int jj = 0; jj < 3; ++jj) fa[kk][ii][jj] = 0.0f;
      const float* g32 = a.asmtab32 + (size_t)e * kGP * kAsmStride32;
      const float c5f = (float)c5, c6f = (float)c6;
#pragma unroll 1
      for (int g = 0; g < NG; ++g) {
        const double* gd = gp + (size_t)g * kAsmStride;
        const float* gf = g32 + (size_t)g * kAsmStride32;
        const float r1 = (float)(gd[0] - xp), r2 = (float)(gd[1] - yp), r3 = (float)(gd[2] - zp);
        const float ir = __builtin_amdgcn_rsqf(r1 * r1 + r2 * r2 + r3 * r3);
        const float eta[3] = {gf[0], gf[1], gf[2]};
        const float rd[3] = {r1 * ir, r2 * ir, r3 * ir};
        const float rdn = (r1 * eta[0] + r2 * eta[1] + r3 * eta[2]) * ir;
        const float c5r2 = c5f * (ir * ir);
        float t[3][3];
#pragma unroll
        for (int ii = 0; ii < 3; ++ii)
#pragma unroll
          for (int jj = 0; jj < 3; ++jj) {
            const float dl = (ii == jj) ? c6f : 0.0f;
            t[jj][ii] = c5r2 * (rdn * (dl + 3.0f * rd[ii] * rd[jj]) + c6f * (eta[jj] * rd[ii] - eta[ii] * rd[jj]));
          }
#pragma unroll
        for (int kk = 0; kk < NC; ++kk) {
          const float p = gf[3 + kk];
#pragma unroll
          for (int ii = 0; ii < 3; ++ii)
#pragma unroll
            for (int jj = 0; jj < 3; ++jj) fa[kk][ii][jj] += t[jj][ii] * p;
        }
      }
#pragma unroll
      for (int kk = 0; kk < NC; ++kk)
#pragma unroll
        for (int ii = 0; ii < 3; ++ii)
#pragma unroll
          for (int jj = 0; jj < 3; ++jj) acc[kk][ii][jj] = (double)fa[kk][ii][jj];
    } else {
    // The traction kernel is summed as its symmetric part
    //   S(i,j) = c5/r^2 (dr/dn) (c6 d_ij + 3 r_,i r_,j)       (6 entries)
    // and its antisymmetric part
    //   W(i,j) = c5/r^2 c6 (n_j r_,i - n_i r_,j)               (3 entries)
    // per shape function; h(j, 3k+i) = S(i,j) + W(i,j) is formed once per
    // element.
    double aS[NC][6], aW[NC][3];  // S: 00 11 22 01 12 02; W: 01 12 02
#pragma unroll
    for (int kk = 0; kk < NC; ++kk) {
#pragma unroll
      for (int c = 0; c < 6; ++c) aS[kk][c] = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) aW[kk][c] = 0.0;
    }
    // a Gauss point's data (position, normal, weights) are loaded one point
    // ahead, in flight during the previous point's arithmetic
    double cur[6 + NC];
#pragma unroll
    for (int q = 0; q < 6 + NC; ++q) cur[q] = gp[q];
#pragma unroll 1
    for (int g = 0; g < NG; ++g) {
      double nxt[6 + NC];
      const double* gq = gp + (size_t)(g + 1 < NG ? g + 1 : g) * kAsmStride;
#pragma unroll
      for (int q = 0; q < 6 + NC; ++q) nxt[q] = gq[q];
      const double r1 = cur[0] - xp, r2 = cur[1] - yp, r3 = cur[2] - zp;
      const double ir = inv_sqrt(r1 * r1 + r2 * r2 + r3 * r3);
      const double d0 = r1 * ir, d1v = r2 * ir, d2 = r3 * ir;
      const double rdn = d0 * cur[3] + d1v * cur[4] + d2 * cur[5];
      const double c5r2 = c5 * (ir * ir);
      const double sc = c5r2 * rdn, wc = c5r2 * c6;
      const double s3 = 3.0 * sc, sd = sc * c6;
      const double q0 = s3 * d0, q1 = s3 * d1v, q2 = s3 * d2;
      const double S[6] = {q0 * d0 + sd, q1 * d1v + sd, q2 * d2 + sd, q0 * d1v, q1 * d2, q0 * d2};
      const double u0 = wc * d0, u1 = wc * d1v, u2 = wc * d2;
      const double W[3] = {cur[4] * u0 - cur[3] * u1, cur[5] * u1 - cur[4] * u2, cur[5] * u0 - cur[3] * u2};
#pragma unroll
      for (int kk = 0; kk < NC; ++kk) {
        const double p = cur[6 + kk];
#pragma unroll
        for (int c = 0; c < 6; ++c) aS[kk][c] += S[c] * p;
#pragma unroll
        for (int c = 0; c < 3; ++c) aW[kk][c] += W[c] * p;
      }
#pragma unroll
      for (int q = 0; q < 6 + NC; ++q) cur[q] = nxt[q];
    }
#pragma unroll
    for (int kk = 0; kk < NC; ++kk) {
      acc[kk][0][0] = aS[kk][0];
      acc[kk][1][1] = aS[kk][1];
      acc[kk][2][2] = aS[kk][2];
      acc[kk][0][1] = aS[kk][3] + aW[kk][0];
      acc[kk][1][0] = aS[kk][3] - aW[kk][0];
      acc[kk][1][2] = aS[kk][4] + aW[kk][1];
      acc[kk][2][1] = aS[kk][4] - aW[kk][1];
      acc[kk][0][2] = aS[kk][5] + aW[kk][2];
      acc[kk][2][0] = aS[kk][5] - aW[kk][2];
    }
    }
    const int kb = inf->kode, k0 = inf->k0mask;
    if (valid && !sing) {
      const double* ebc = ebc_all + (size_t)e * 27;
      const double* evp = evp_all + (size_t)e * 27;
#pragma unroll
      for (int k = 0; k < NC; ++k)
#pragma unroll
        for (int ld = 0; ld < 3; ++ld) {
          const int bit = 3 * k + ld;
          const double h0 = acc[k][ld][0], h1 = acc[k][ld][1], h2 = acc[k][ld][2];
          if (((kb >> ld) & 1) == 0) {
            const double b = ebc[bit];
            fi0 -= h0 * b;
            fi1 -= h1 * b;
            fi2 -= h2 * b;
          } else if ((k0 >> bit) & 1) {
            const double v = evp[bit];
            fi0 -= h0 * v;
            fi1 -= h1 * v;
            fi2 -= h2 * v;
          }
          d1[0 + 3 * ld] -= h0;
          d1[1 + 3 * ld] -= h1;
          d1[2 + 3 * ld] -= h2;
        }
    }
    if (a.write_A && valid) {
      // column blocks of A by the write plan: a chain's first element reads
      // A (or starts from zero), its last stores, the ones between carry the
      // block in LDS.  A lane whose node is in the element adds nothing but
      // still passes its block along.  Reads of A are issued one node ahead.
      const unsigned* eop = a.elop + (size_t)e * kElemOps;
      auto load_cols = [&](int k, double (&o)[3][3]) {
        const unsigned op = eop[k];
        const int rd = (op & kOpFresh) ? (op_wcols(op) & ~op_zcols(op)) : 0;
        const int ik = inf->node[k];
#pragma unroll
        for (int ld = 0; ld < 3; ++ld) {
          const double* col = Aout + (long long)(3 * ik + ld) * a.lda + 3LL * in;
#pragma unroll
          for (int q = 0; q < 3; ++q) o[ld][q] = ((rd >> ld) & 1) ? col[q] : 0.0;
        }
      };
      double nxt[3][3];
      load_cols(0, nxt);
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        const unsigned op = eop[k];
        const int ik = inf->node[k];
        double v[3][3];
#pragma unroll
        for (int ld = 0; ld < 3; ++ld)
#pragma unroll
          for (int q = 0; q < 3; ++q) v[ld][q] = nxt[ld][q];
        if (k + 1 < NC) load_cols(k + 1, nxt);
        if (!(op & kOpFresh)) {
          const int s = op_slot_in(op) & (kCarrySlots - 1);
#pragma unroll
          for (int ld = 0; ld < 3; ++ld)
#pragma unroll
            for (int q = 0; q < 3; ++q) v[ld][q] = carry[s][3 * ld + q][threadIdx.x];
        }
        if (!sing) {
#pragma unroll
          for (int ld = 0; ld < 3; ++ld)
            if (((kb >> ld) & 1) && !((k0 >> (3 * k + ld)) & 1)) {
#pragma unroll
              for (int q = 0; q < 3; ++q) v[ld][q] += acc[k][ld][q];
            }
        }
        if (op & kOpLast) {
          const int wc = op_wcols(op);
#pragma unroll
          for (int ld = 0; ld < 3; ++ld)
            if ((wc >> ld) & 1) {
              double* col = Aout + (long long)(3 * ik + ld) * a.lda + 3LL * in;
#pragma unroll
              for (int q = 0; q < 3; ++q) col[q] = v[ld][q];
            }
        } else {
          const int s = op_slot_out(op) & (kCarrySlots - 1);
#pragma unroll
          for (int ld = 0; ld < 3; ++ld)
#pragma unroll
            for (int q = 0; q < 3; ++q) carry[s][3 * ld + q][threadIdx.x] = v[ld][q];
        }
      }
    }
  }
  if (valid) {
    const size_t o = (size_t)chunk * a.n + in;
#pragma unroll
    for (int i = 0; i < 9; ++i) a.dpart[o * 9 + i] = d1[i];
    double* f = a.fpart + o * 3;
    f[0] = fi0;
    f[1] = fi1;
    f[2] = fi2;
  }
}

// G pass: g_elem of extin8 with funkelvin's displacement kernel, then the
// G terms of ghmat8 (assembly_mod.f90:131-133, 164-166, 179-185).  FAR32 as
// in the H pass.
template <int NC, bool FAR32>
__global__ __launch_bounds__(64) void k_gpass(AsmArgs a, const int* __restrict__ gchunks, int coff,
                                               const double* __restrict__ asmtab, const ElemInfo* __restrict__ einfo,
                                               const double* __restrict__ ebc_all, double* __restrict__ Aout) {
  constexpr int NG = (NC == 9) ? 36 : 16;
  const int gidx = coff + blockIdx.y;
  const int chunk = gchunks[gidx];
  const int e0 = a.chunk_start[chunk], e1 = a.chunk_start[chunk + 1];
  const int in = blockIdx.x * 64 + threadIdx.x;
  const bool valid = in < a.n;
  const int inr = valid ? in : a.n - 1;
  const double xp = a.x[inr], yp = a.y[inr], zp = a.z[inr];
  const double c3 = a.c3, c4 = a.c4, mu = a.mu;
  double fi0 = 0.0, fi1 = 0.0, fi2 = 0.0;
  for (int e = e0; e < e1; ++e) {
    const ElemInfo* inf = einfo + e;
    if (!inf->needG) continue;
    bool sing = false;
#pragma unroll
    for (int k = 0; k < NC; ++k) sing |= (inf->node[k] == in);
    double acc[NC][6];  // symmetric U: 00 11 22 01 12 02
#pragma unroll
    for (int kk = 0; kk < NC; ++kk)
#pragma unroll
      for (int c = 0; c < 6; ++c) acc[kk][c] = 0.0;
    const double* gp = asmtab + (size_t)e * kGP * kAsmStride;
    if (FAR32 && is_far(a.elfar, e, xp, yp, zp)) {
      float fa[NC][6];
#pragma unroll
      for (int kk = 0; kk < NC; ++kk)
#pragma unroll
        for (int c = 0; c < 6; ++c) fa[kk][c] = 0.0f;
      const float* g32 = a.asmtab32 + (size_t)e * kGP * kAsmStride32;
      const float c3f = (float)c3, c4f = (float)c4;
#pragma unroll 2
      for (int g = 0; g < NG; ++g) {
        const double* gd = gp + (size_t)g * kAsmStride;
        const float* gf = g32 + (size_t)g * kAsmStride32;
        const float r1 = (float)(gd[0] - xp), r2 = (float)(gd[1] - yp), r3 = (float)(gd[2] - zp);
        const float ir = __builtin_amdgcn_rsqf(r1 * r1 + r2 * r2 + r3 * r3);
        const float rd0 = r1 * ir, rd1 = r2 * ir, rd2 = r3 * ir;
        const float c3r = c3f * ir;
        const float u[6] = {(c4f + rd0 * rd0) * c3r, (c4f + rd1 * rd1) * c3r, (c4f + rd2 * rd2) * c3r,
                            (rd0 * rd1) * c3r,       (rd1 * rd2) * c3r,       (rd0 * rd2) * c3r};
#pragma unroll
        for (int kk = 0; kk < NC; ++kk) {
          const float p = gf[3 + kk];
#pragma unroll
          for (int c = 0; c < 6; ++c) fa[kk][c] += u[c] * p;
        }
      }
#pragma unroll
      for (int kk = 0; kk < NC; ++kk)
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[kk][c] = (double)fa[kk][c];
    } else
#pragma unroll 2
    for (int g = 0; g < NG; ++g, gp += kAsmStride) {
      const double r1 = gp[0] - xp, r2 = gp[1] - yp, r3 = gp[2] - zp;
      const double ir = inv_sqrt(r1 * r1 + r2 * r2 + r3 * r3);
      const double rd0 = r1 * ir, rd1 = r2 * ir, rd2 = r3 * ir;
      const double c3r = c3 * ir;
      const double u[6] = {(c4 + rd0 * rd0) * c3r, (c4 + rd1 * rd1) * c3r, (c4 + rd2 * rd2) * c3r,
                           (rd0 * rd1) * c3r,      (rd1 * rd2) * c3r,      (rd0 * rd2) * c3r};
#pragma unroll
      for (int kk = 0; kk < NC; ++kk) {
        const double p = gp[6 + kk];
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[kk][c] += u[c] * p;
      }
    }
    if (valid) {
      const int kb = inf->kode;
      const double* ebc = ebc_all + (size_t)e * 27;
      // first writes of a column (write plan bit 3k+ld) start from zero; a
      // lane whose node is in the element adds nothing, and stores only
      // those first writes
      const unsigned gz = a.write_A ? a.elop[(size_t)e * kElemOps + 9] : 0u;
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        const int ik = inf->node[k];
        // the node's A columns: all loads in flight before the updates
        double old[3][3] = {};
        if (a.write_A && !sing && (kb & 7) != 7) {
#pragma unroll
          for (int ld = 0; ld < 3; ++ld) {
            if (((kb >> ld) & 1) || ((gz >> (3 * k + ld)) & 1)) continue;
            const double* col = Aout + (long long)(3 * ik + ld) * a.lda + 3LL * in;
#pragma unroll
            for (int q = 0; q < 3; ++q) old[ld][q] = col[q];
          }
        }
#pragma unroll
        for (int ld = 0; ld < 3; ++ld) {
          // g(jj, 3k+ld) = U(jj, ld)
          const int s0 = (ld == 0) ? 0 : (ld == 1 ? 3 : 5);
          const int s1 = (ld == 0) ? 3 : (ld == 1 ? 1 : 4);
          const int s2 = (ld == 0) ? 5 : (ld == 1 ? 4 : 2);
          const double g0 = acc[k][s0], g1 = acc[k][s1], g2 = acc[k][s2];
          if (((kb >> ld) & 1) == 0) {
            if (a.write_A) {
              double* col = Aout + (long long)(3 * ik + ld) * a.lda + 3LL * in;
              if (!sing) {
                col[0] = old[ld][0] - g0 * mu;
                col[1] = old[ld][1] - g1 * mu;
                col[2] = old[ld][2] - g2 * mu;
              } else if ((gz >> (3 * k + ld)) & 1) {
                col[0] = 0.0;
                col[1] = 0.0;
                col[2] = 0.0;
              }
            }
          } else if (!sing) {
            const double b = ebc[3 * k + ld];
            fi0 += g0 * b;
            fi1 += g1 * b;
            fi2 += g2 * b;
          }
        }
      }
    }
  }
  if (valid) {
    double* f = a.fpart + ((size_t)gidx * a.n + in) * 3;
    f[0] = fi0;
    f[1] = fi1;
    f[2] = fi2;
  }
}

// Singular pairs, locin8 + triloc8 (assembly_mod.f90:386-616): one 64-lane
// workgroup per pair; lanes evaluate Gauss points of the Duffy-split element
// into LDS, then each lane sums whole output entries over the points in the
// reference's order (sub-triangle, i, j).
constexpr int kMaxSingGP = 4 * 36;

__global__ __launch_bounds__(64) void k_singular(int npairs, const SingPair* __restrict__ pairs,
                                                 const int* __restrict__ nconec,
                                                 const int* __restrict__ kconec0,
                                                 const double* __restrict__ x,
                                                 const double* __restrict__ y,
                                                 const double* __restrict__ z,
                                                 const SubTri* __restrict__ subtri, Material mat,
                                                 double* __restrict__ out) {
  __shared__ double sK[kMaxSingGP][18];  // u(jj,ii) at [jj+3ii], t at 9+[jj+3ii]
  __shared__ double sP[kMaxSingGP][9];
  const int p = blockIdx.x;
  if (p >= npairs) return;
  const SingPair sp = pairs[p];
  const int e = sp.elem, nc = nconec[e];
  const int kind = (nc == 9) ? 0 : 1;
  double xj1[9], xj2[9], xj3[9];
  for (int m = 0; m < 9; ++m) {
    int nd = m < nc ? kconec0[e * 9 + m] : 0;
    xj1[m] = m < nc ? x[nd] : 0.0;
    xj2[m] = m < nc ? y[nd] : 0.0;
    xj3[m] = m < nc ? z[nd] : 0.0;
  }
  const double xi0 = xj1[sp.nodo - 1], xi1 = xj2[sp.nodo - 1], xi2 = xj3[sp.nodo - 1];
  const int ngp = sp.nsub * 36;
  for (int g = threadIdx.x; g < ngp; g += 64) {
    const SubTri st = subtri[(kind * 10 + sp.nodo) * 4 + g / 36];
    const int i = (g % 36) / 6, j = g % 6;
    const double ss1 = 0.5 * (c_gi[i] + 1.0), ss2 = 0.5 * (c_gi[j] + 1.0);
    const double g1 = (1.0 - ss1) * st.v1[0] + ss1 * (1.0 - ss2) * st.v1[1] + ss1 * ss2 * st.v1[2];
    const double g2 = (1.0 - ss1) * st.v2[0] + ss1 * (1.0 - ss2) * st.v2[1] + ss1 * ss2 * st.v2[2];
    double f[9], fd1[9], fd2[9];
    shape(nc, g1, g2, f, fd1, fd2);
    double s1, s2, s3;
    normal_dc(nc, fd1, fd2, xj1, xj2, xj3, s1, s2, s3);
    const double xja = sqrt(s1 * s1 + s2 * s2 + s3 * s3);
    const double eta[3] = {s1 / xja, s2 / xja, s3 / xja};
    double xq0 = 0.0, xq1 = 0.0, xq2 = 0.0;
    for (int m = 0; m < nc; ++m) {
      xq0 += f[m] * xj1[m];
      xq1 += f[m] * xj2[m];
      xq2 += f[m] * xj3[m];
    }
    for (int m = 0; m < 9; ++m)
      sP[g][m] = (m < nc) ? f[m] * xja * (c_gi[i] + 1.0) * st.area * 0.25 * c_ome[i] * c_ome[j] : 0.0;
    // funkelvin, assembly_mod.f90:618-650
    const double r1 = xq0 - xi0, r2 = xq1 - xi1, r3 = xq2 - xi2;
    const double r = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
    const double rdn = (r1 * eta[0] + r2 * eta[1] + r3 * eta[2]) / r;
    const double rd[3] = {r1 / r, r2 / r, r3 / r};
    const double c3r = mat.c3 / r, c5r2 = mat.c5 / (r * r);
    for (int ii = 0; ii < 3; ++ii)
      for (int jj = 0; jj < 3; ++jj) {
        const double dl = (ii == jj) ? 1.0 : 0.0;
        sK[g][jj + 3 * ii] = (mat.c4 * dl + rd[ii] * rd[jj]) * c3r;
        sK[g][9 + jj + 3 * ii] =
            c5r2 * (rdn * (mat.c6 * dl + 3.0 * rd[ii] * rd[jj]) + mat.c6 * (eta[jj] * rd[ii] - eta[ii] * rd[jj]));
      }
  }
  __syncthreads();
  // outputs: h(jj, ind) at [jj + 3 ind], g at 81 + [jj + 3 ind]
  for (int o = threadIdx.x; o < 162; o += 64) {
    const int isg = o >= 81;
    const int q = isg ? o - 81 : o;
    const int jj = q % 3, ind = q / 3, kk = ind / 3, ii = ind % 3;
    double s = 0.0;
    if (kk < nc) {
      const int c = (isg ? 0 : 9) + jj + 3 * ii;
      for (int g = 0; g < ngp; ++g) s += sK[g][c] * sP[g][kk];
    }
    out[(size_t)p * 162 + o] = s;
  }
}

struct FinArgs {
  int n, nchunk, ngchunk;
  long long lda;
  const double* dpart;
  const double* fpart_h;
  const double* fpart_g;
  const int* pair_ptr;
  const SingPair* pairs;
  const double* sing;
  const ElemInfo* einfo;
  const double* ebc;
  const double* evp;
  const int* node_k0;
  const double* node_vp;
  double mu;
  double* A;
  double* FI;
  int write_A;
};

// Per collocation node: reduce the chunk partials, apply the singular
// elements exactly as ghmat8's nodo>0 branch (assembly_mod.f90:112-190), route
// the regular rigid-body sums (192-218), store the diagonal block and FI.
__global__ void k_finalize(FinArgs a) {
  const int in = blockIdx.x * blockDim.x + threadIdx.x;
  if (in >= a.n) return;
  double D[9], F[3] = {0.0, 0.0, 0.0}, Dg[9];
  for (int i = 0; i < 9; ++i) D[i] = Dg[i] = 0.0;
  for (int c = 0; c < a.nchunk; ++c) {
    const double* dp = a.dpart + ((size_t)c * a.n + in) * 9;
    for (int i = 0; i < 9; ++i) D[i] += dp[i];
    const double* fp = a.fpart_h + ((size_t)c * a.n + in) * 3;
    F[0] += fp[0];
    F[1] += fp[1];
    F[2] += fp[2];
  }
  for (int c = 0; c < a.ngchunk; ++c) {
    const double* fp = a.fpart_g + ((size_t)c * a.n + in) * 3;
    F[0] += fp[0];
    F[1] += fp[1];
    F[2] += fp[2];
  }
  const double mu = a.mu;
  for (int p = a.pair_ptr[in]; p < a.pair_ptr[in + 1]; ++p) {
    const SingPair sp = a.pairs[p];
    const int e = sp.elem;
    const ElemInfo inf = a.einfo[e];
    const double* hs = a.sing + (size_t)p * 162;
    const double* gs = hs + 81;
    const int nodo = sp.nodo - 1;
    double d2[9];
    for (int ld = 0; ld < 3; ++ld)
      for (int jj = 0; jj < 3; ++jj) {
        double s = 0.0;
        for (int k = 0; k < inf.nconec; ++k)
          if (k != nodo) s -= hs[jj + 3 * (3 * k + ld)];
        d2[jj + 3 * ld] = s;
      }
    for (int ld = 0; ld < 3; ++ld) {
      const bool traction = (inf.kode >> ld) & 1;
      for (int k = 0; k < inf.nconec; ++k) {
        const int ik = inf.node[k];
        const int ind = 3 * k + ld;
        const double b = a.ebc[(size_t)e * 27 + ind];
        double* col = a.A + (long long)(3 * ik + ld) * a.lda + 3LL * in;
        if (!traction) {
          for (int jj = 0; jj < 3; ++jj) {
            const double gv = -gs[jj + 3 * ind] * mu;
            if (ik == in) Dg[jj + 3 * ld] += gv;
            else if (a.write_A) col[jj] += gv;
            F[jj] -= ((ik != in) ? hs[jj + 3 * ind] : d2[jj + 3 * ld]) * b;
          }
        } else if (((inf.k0mask >> ind) & 1) == 0) {
          for (int jj = 0; jj < 3; ++jj) {
            F[jj] += gs[jj + 3 * ind] * b;
            if (ik == in) Dg[jj + 3 * ld] += d2[jj + 3 * ld];
            else if (a.write_A) col[jj] += hs[jj + 3 * ind];
          }
        } else {
          const double v = a.evp[(size_t)e * 27 + ind];
          for (int jj = 0; jj < 3; ++jj)
            F[jj] += gs[jj + 3 * ind] * b - ((ik != in) ? hs[jj + 3 * ind] : d2[jj + 3 * ld]) * v;
        }
      }
    }
  }
  for (int ld = 0; ld < 3; ++ld) {
    if (a.node_k0[3 * in + ld] == 0) {
      for (int jj = 0; jj < 3; ++jj) Dg[jj + 3 * ld] += D[jj + 3 * ld];
    } else {
      const double v = a.node_vp[3 * in + ld];
      for (int jj = 0; jj < 3; ++jj) F[jj] -= D[jj + 3 * ld] * v;
    }
  }
  if (a.write_A)
    for (int ld = 0; ld < 3; ++ld)
      for (int jj = 0; jj < 3; ++jj) a.A[(long long)(3 * in + ld) * a.lda + 3LL * in + jj] = Dg[jj + 3 * ld];
  for (int jj = 0; jj < 3; ++jj) a.FI[3LL * in + jj] = F[jj];
}

// The entries of A that no assembly pass writes: rows [N, lda) of every
// column (k_zero_tail), and the columns of the write plan's unwritten list
// (k_zero_cols, one workgroup per column).
__global__ void k_zero_tail(long long N, long long lda, double* __restrict__ A) {
  const long long pad = lda - N;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= pad * N) return;
  A[(t / pad) * lda + N + t % pad] = 0.0;
}
__global__ void k_zero_cols(const int* __restrict__ cols, long long lda, double* __restrict__ A) {
  double* col = A + (long long)cols[blockIdx.x] * lda;
  for (long long r = threadIdx.x; r < lda; r += blockDim.x) col[r] = 0.0;
}

int geometry_run(qd_context* h) {
  const size_t slots = (size_t)h->ne * kGP;
  QD_ALLOC(h, h->geo, slots * 16);
  QD_ALLOC(h, h->asmtab, slots * kAsmStride);
  QD_ALLOC(h, h->asmtab32, slots * kAsmStride32);
  QD_ALLOC(h, h->elfar, 4 * (size_t)h->ne);
  const int bs = 256;
  const int nb = (int)((slots + bs - 1) / bs);
  hipLaunchKernelGGL(k_geometry, dim3(nb), dim3(bs), 0, h->stream, h->ne, h->nconec.p, h->kconec0.p,
                     h->x.p, h->y.p, h->z.p, h->geo.p, h->asmtab.p, h->asmtab32.p);
  hipLaunchKernelGGL(k_elem_far, dim3((h->ne + 255) / 256), dim3(256), 0, h->stream, h->ne, h->nconec.p,
                     h->kconec0.p, h->x.p, h->y.p, h->z.p, h->elfar.p);
  QD_HIP(h, hipGetLastError());
  return QD_OK;
}

int assembly_run(qd_context* h, double* fi_out_dev, bool write_A, bool far32, bool into_pre) {
  const int n = h->n;
  const long long N = 3LL * n;
  const int nchunk = (int)h->h_chunk_start.size() - 1;
  const int ngchunk = (int)h->h_gchunks.size();
  DevBuf<double>& Abuf = into_pre ? h->A_pre : h->A;
  DevBuf<double>& FIbuf = into_pre ? h->FI_pre : h->FI;
  if (write_A) {
    long long lda = (N + 63) / 64 * 64;  // 512-byte aligned columns
    h->lda = lda;
    QD_ALLOC(h, Abuf, (size_t)lda * (size_t)N);
    QD_ALLOC(h, FIbuf, (size_t)N);
    // every other entry is written by the passes below (the write plan)
    const long long tail = (lda - N) * N;
    if (tail > 0)
      hipLaunchKernelGGL(k_zero_tail, dim3((unsigned)((tail + 255) / 256)), dim3(256), 0, h->stream, N, lda,
                         Abuf.p);
    if (!h->h_unwritten.empty())
      hipLaunchKernelGGL(k_zero_cols, dim3((unsigned)h->h_unwritten.size()), dim3(256), 0, h->stream,
                         h->unwritten.p, lda, Abuf.p);
    QD_HIP(h, hipGetLastError());
  }
  QD_ALLOC(h, h->dpart, (size_t)nchunk * n * 9);
  QD_ALLOC(h, h->fpart_h, (size_t)nchunk * n * 3);
  QD_ALLOC(h, h->fpart_g, std::max<size_t>((size_t)ngchunk * n * 3, 1));
  QD_ALLOC(h, h->sing_hg, h->h_pairs.size() * 162);

  AsmArgs a{};
  a.n = n;
  a.ne = h->ne;
  a.lda = h->lda;
  a.x = h->x.p;
  a.y = h->y.p;
  a.z = h->z.p;
  a.asmtab = h->asmtab.p;
  a.einfo = h->einfo.p;
  a.ebc = h->ebc.p;
  a.evp = h->evp.p;
  a.chunk_start = h->chunk_start.p;
  a.A = Abuf.p;
  a.dpart = h->dpart.p;
  a.asmtab32 = h->asmtab32.p;
  a.elfar = h->elfar.p;
  a.elop = h->elop.p;
  a.mu = h->mat.mu;
  a.c3 = h->mat.c3;
  a.c4 = h->mat.c4;
  a.c5 = h->mat.c5;
  a.c6 = h->mat.c6;
  a.write_A = write_A ? 1 : 0;
  const int tiles = (n + 63) / 64;
  const int ncol = (int)h->h_colour_start.size() - 1;
  // chunks of one colour hold one element kind only if the mesh is sorted;
  // launch per (colour, kind) by splitting on the first element's kind.
  for (int col = 0; col < ncol; ++col) {
    for (int pass = 0; pass < 2; ++pass) {
      const std::vector<int>& list = pass == 0 ? h->h_colour_chunks : h->h_gchunks;
      const std::vector<int>& start = pass == 0 ? h->h_colour_start : h->h_gcolour_start;
      int b = start[col], e = start[col + 1];
      // runs of same-kind chunks within the colour
      int i = b;
      while (i < e) {
        int kind = h->hnconec[h->h_chunk_start[list[i]]];
        int j = i;
        while (j < e && h->hnconec[h->h_chunk_start[list[j]]] == kind) ++j;
        dim3 grid(tiles, j - i);
        // work: Gauss-point evaluations x FP64 flops of one evaluation as
        // these passes compute it (H: traction kernel with 1/r ~56 flops + 18
        // per shape function of accumulation; G: displacement kernel ~40 + 12
        // per shape function)
        double gpe = 0.0;
        for (int q = i; q < j; ++q)
          for (int el = h->h_chunk_start[list[q]]; el < h->h_chunk_start[list[q] + 1]; ++el)
            gpe += (double)n * (kind == 9 ? 36.0 : 16.0);
        const double fpg = pass == 0 ? 56.0 + 18.0 * kind : 40.0 + 12.0 * kind;
        ProfScope ps(h, pass == 0 ? QD_PROF_ASSEMBLY_H : QD_PROF_ASSEMBLY_G, gpe * fpg);
        if (pass == 0) {
          a.fpart = h->fpart_h.p;
          auto kh = kind == 9 ? (far32 ? k_hpass<9, true> : k_hpass<9, false>)
                              : (far32 ? k_hpass<6, true> : k_hpass<6, false>);
          hipLaunchKernelGGL(kh, grid, dim3(64), 0, h->stream, a, h->colour_chunks.p, i, h->asmtab.p, h->einfo.p,
                             h->ebc.p, h->evp.p, Abuf.p);
        } else {
          a.fpart = h->fpart_g.p;
          auto kg = kind == 9 ? (far32 ? k_gpass<9, true> : k_gpass<9, false>)
                              : (far32 ? k_gpass<6, true> : k_gpass<6, false>);
          hipLaunchKernelGGL(kg, grid, dim3(64), 0, h->stream, a, h->gchunks.p, i, h->asmtab.p, h->einfo.p,
                             h->ebc.p, Abuf.p);
        }
        QD_HIP(h, hipGetLastError());
        i = j;
      }
    }
  }
  const int npairs = (int)h->h_pairs.size();
  hipLaunchKernelGGL(k_singular, dim3(npairs), dim3(64), 0, h->stream, npairs, h->pairs.p, h->nconec.p,
                     h->kconec0.p, h->x.p, h->y.p, h->z.p, h->subtri.p, h->mat, h->sing_hg.p);
  QD_HIP(h, hipGetLastError());
  FinArgs f{};
  f.n = n;
  f.nchunk = nchunk;
  f.ngchunk = ngchunk;
  f.lda = h->lda;
  f.dpart = h->dpart.p;
  f.fpart_h = h->fpart_h.p;
  f.fpart_g = h->fpart_g.p;
  f.pair_ptr = h->pair_ptr.p;
  f.pairs = h->pairs.p;
  f.sing = h->sing_hg.p;
  f.einfo = h->einfo.p;
  f.ebc = h->ebc.p;
  f.evp = h->evp.p;
  f.node_k0 = h->node_k0.p;
  f.node_vp = h->node_vp.p;
  f.mu = h->mat.mu;
  f.A = Abuf.p;
  f.FI = fi_out_dev ? fi_out_dev : FIbuf.p;
  f.write_A = write_A ? 1 : 0;
  hipLaunchKernelGGL(k_finalize, dim3((n + 127) / 128), dim3(128), 0, h->stream, f);
  QD_HIP(h, hipGetLastError());
  return QD_OK;
}

}  // namespace qd
