// qd_somigliana.hip -- internal-point displacements and stresses
// (compute_internal_points, stress_mod.f90:511-781) and the boundary stress
// recovery (compute_boundary_stresses, stress_mod.f90:366-406) on gfx950.
//
//  * Gauss tables: a far table (the 6x6 / 4x4 element rules, pre-weighted
//    with boundary u and t, stress_mod.f90:580-651) and a near table (40x40
//    per element, 654-726), 12 doubles per point: x(3) n(3) u*w(3) t*w(3).
//  * far pass: one lane per internal point, 64 points per wave, a slice of
//    elements per wave; the element's far points are wave-uniform and come
//    through scalar loads.  Lanes for which the element is near skip it.
//  * near pass: one wave per internal point; the wave finds the point's near
//    elements with a ballot, spreads each element's 1600 near points over the
//    lanes and reduces the 9 sums across the wave.
//  * final: per point, slices then near sums, then principal stresses and
//    von Mises (stress_mod.f90:299-359).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "qd_internal.h"
#include "qd_device.h"

namespace qd {

constexpr int kSomiStride = 12;
// far table rows add the products that do not depend on the point:
// n.U and the symmetrised n (x) U (xx yy zz as 2 n_i U_i; xy yz zx)
constexpr int kFarStride = 20;
constexpr int kNear = kNearGL * kNearGL;

__constant__ double c_gl40x[kNearGL];
__constant__ double c_gl40w[kNearGL];
__constant__ double c_ome6[6] = {0.467913934572691, 0.467913934572691, 0.360761573048139,
                                 0.360761573048139, 0.171324492379170, 0.171324492379170};
__constant__ double c_omei4[4] = {0.1739274226, 0.3260725774, 0.3260725774, 0.1739274226};
__constant__ double c_omej4[4] = {0.1355069134, 0.2034645680, 0.1298475476, 0.0311809709};

// gauss_legendre_1d (stress_mod.f90:475-503), evaluated on the host.
void gauss_legendre(int npts, double* xg, double* wg) {
  const double pi = 3.14159265358979323846;
  int m_half = (npts + 1) / 2;
  for (int i = 1; i <= m_half; ++i) {
    double z = cos(pi * ((double)i - 0.25) / ((double)npts + 0.5));
    double z1, p1 = 0.0, p2 = 0.0, p3, pp = 0.0;
    for (int iter = 1; iter <= 100; ++iter) {
      p1 = 1.0;
      p2 = 0.0;
      for (int j = 1; j <= npts; ++j) {
        p3 = p2;
        p2 = p1;
        p1 = ((2.0 * (double)j - 1.0) * z * p2 - (double)(j - 1) * p3) / (double)j;
      }
      pp = (double)npts * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
      if (fabs(z - z1) <= 1.0e-15) break;
    }
    xg[i - 1] = -z;
    xg[npts - i] = z;
    wg[i - 1] = 2.0 / ((1.0 - z * z) * pp * pp);
    wg[npts - i] = wg[i - 1];
  }
}

// Element centroid and near threshold, stress_mod.f90:584-604.
__global__ void k_centres(int ne, const int* __restrict__ nconec, const int* __restrict__ kconec0,
                          const double* __restrict__ x, const double* __restrict__ y,
                          const double* __restrict__ z, double* __restrict__ cen) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int nc = nconec[e];
  double cx = 0.0, cy = 0.0, cz = 0.0;
  for (int m = 0; m < nc; ++m) {
    int nd = kconec0[e * 9 + m];
    cx += x[nd];
    cy += y[nd];
    cz += z[nd];
  }
  cx /= (double)nc;
  cy /= (double)nc;
  cz /= (double)nc;
  double dmax2 = 0.0;
  for (int m = 0; m < nc; ++m) {
    int nd = kconec0[e * 9 + m];
    double dx = x[nd] - cx, dy = y[nd] - cy, dz = z[nd] - cz;
    dmax2 = fmax(dmax2, dx * dx + dy * dy + dz * dz);
  }
  cen[4 * e + 0] = cx;
  cen[4 * e + 1] = cy;
  cen[4 * e + 2] = cz;
  cen[4 * e + 3] = (3.5 * 3.5) * dmax2;
}

// Far table, stress_mod.f90:606-650.
__global__ void k_far_table(int ne, const int* __restrict__ nconec, const int* __restrict__ kconec0,
                            const int* __restrict__ far_start, const double* __restrict__ geo,
                            const double* __restrict__ fi, const double* __restrict__ dfi,
                            double* __restrict__ tab) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= (long long)ne * kGP) return;
  const int e = (int)(tid / kGP), k = (int)(tid % kGP);
  const int nc = nconec[e];
  const int npg = nc == 9 ? 6 : 4;
  if (k >= npg * npg) return;
  const int i = k / npg, j = k % npg;
  const size_t blk = (size_t)ne * kGP, slot = (size_t)e * kGP + k;
  const double xja = geo[3 * blk + slot];
  const double wg = (nc == 9) ? c_ome6[i] * c_ome6[j] * xja : c_omei4[i] * c_omej4[j] * xja;
  double ug[3] = {0.0, 0.0, 0.0}, tg[3] = {0.0, 0.0, 0.0};
  for (int m = 0; m < nc; ++m) {
    const int ik = kconec0[e * 9 + m];
    const double fv = geo[7 * blk + slot * 9 + m];
    for (int d = 0; d < 3; ++d) {
      ug[d] += fv * fi[3 * ik + d];
      tg[d] += fv * dfi[e * 27 + 3 * m + d];
    }
  }
  double* t = tab + (size_t)(far_start[e] + k) * kFarStride;
  t[0] = geo[4 * blk + slot];
  t[1] = geo[5 * blk + slot];
  t[2] = geo[6 * blk + slot];
  t[3] = geo[0 * blk + slot];
  t[4] = geo[1 * blk + slot];
  t[5] = geo[2 * blk + slot];
  for (int d = 0; d < 3; ++d) {
    t[6 + d] = ug[d] * wg;
    t[9 + d] = tg[d] * wg;
  }
  const double* n = t + 3;
  const double* U = t + 6;
  t[12] = n[0] * U[0] + n[1] * U[1] + n[2] * U[2];
  for (int d = 0; d < 3; ++d) {
    t[13 + d] = 2.0 * n[d] * U[d];
    t[16 + d] = n[d] * U[(d + 1) % 3] + n[(d + 1) % 3] * U[d];
  }
  t[19] = 0.0;
}

// Near table (40x40 per element), stress_mod.f90:659-726.
__global__ void k_near_table(int ne, const int* __restrict__ nconec, const int* __restrict__ kconec0,
                             const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, const double* __restrict__ fi,
                             const double* __restrict__ dfi, double* __restrict__ tab) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= (long long)ne * kNear) return;
  const int e = (int)(tid / kNear), g = (int)(tid % kNear);
  const int i = g / kNearGL, j = g % kNearGL;
  const int nc = nconec[e];
  double g1, g2, tri_jac;
  if (nc == 9) {
    g1 = c_gl40x[i];
    g2 = c_gl40x[j];
    tri_jac = 1.0;
  } else {
    g1 = 0.5 * (1.0 + c_gl40x[i]);
    g2 = 0.25 * (1.0 - c_gl40x[i]) * (1.0 + c_gl40x[j]);
    tri_jac = 0.125 * (1.0 - c_gl40x[i]);
  }
  double f[9], fd1[9], fd2[9];
  shape(nc, g1, g2, f, fd1, fd2);
  double xf[3] = {0, 0, 0}, dx1[3] = {0, 0, 0}, dx2[3] = {0, 0, 0}, ug[3] = {0, 0, 0}, tg[3] = {0, 0, 0};
  for (int m = 0; m < nc; ++m) {
    const int ik = kconec0[e * 9 + m];
    const double xm[3] = {x[ik], y[ik], z[ik]};
    for (int d = 0; d < 3; ++d) {
      xf[d] += f[m] * xm[d];
      dx1[d] += fd1[m] * xm[d];
      dx2[d] += fd2[m] * xm[d];
      ug[d] += f[m] * fi[3 * ik + d];
      tg[d] += f[m] * dfi[e * 27 + 3 * m + d];
    }
  }
  const double cv[3] = {dx1[1] * dx2[2] - dx1[2] * dx2[1], dx1[2] * dx2[0] - dx1[0] * dx2[2],
                        dx1[0] * dx2[1] - dx1[1] * dx2[0]};
  const double jac = sqrt(cv[0] * cv[0] + cv[1] * cv[1] + cv[2] * cv[2]);
  const double wg = c_gl40w[i] * c_gl40w[j] * tri_jac * jac;
  double* t = tab + (size_t)tid * kSomiStride;
  for (int d = 0; d < 3; ++d) {
    t[d] = xf[d];
    t[3 + d] = cv[d] / jac;
    t[6 + d] = ug[d] * wg;
    t[9 + d] = tg[d] * wg;
  }
}

struct SomiConst {
  double nu, c4, c6, pr4, cU, cT, cD, cS, sc6;
};

// accumulate_somigliana_range for one Gauss point (stress_mod.f90:807-885);
// sc6 = c6 keeps the reference's displacement T* sign, -c6 the corrected one.
// With r_i the unit vector from the point to the Gauss point, n the normal,
// U = u w and T = t w the weighted boundary displacement and traction, and
// the scalars rdn = r.n, rdt = r.T, rdu = r.U, nuv = n.U, the reference's
// terms are regrouped by the vectors they multiply:
//   u_i   += a_t T_i + a_r r_i + a_u U_i + a_n n_i
//   D_ij - S_ij = C_i r_j + C_j r_i + B_i U_j + B_j U_i + d_ij k_d
// with C_i = k_t T_i + k_u U_i + k_n n_i + k_rr/2 r_i and B_i = k_nu n_i;
// the scalar coefficients are the reference's products of fU = cU/r,
// fT = cT/r^2, fD = cD/r^2 and fS = cS/r^3 with rdn, rdt, rdu and nuv.
// FAR: the row carries n.U and n (x) U (kFarStride layout).
template <bool FAR>
__device__ __forceinline__ void somi_gp(const double* __restrict__ t, double px, double py, double pz,
                                        const SomiConst& k, double* u, double* s, double& sdiag) {
  const double rx = t[0] - px, ry = t[1] - py, rz = t[2] - pz;
  const double inv_r = inv_sqrt(rx * rx + ry * ry + rz * rz);
  const double inv_r2 = inv_r * inv_r;
  const double r[3] = {rx * inv_r, ry * inv_r, rz * inv_r};
  const double n[3] = {t[3], t[4], t[5]};
  const double U[3] = {t[6], t[7], t[8]};
  const double T[3] = {t[9], t[10], t[11]};
  const double rdn = r[0] * n[0] + r[1] * n[1] + r[2] * n[2];
  const double rdt = r[0] * T[0] + r[1] * T[1] + r[2] * T[2];
  const double rdu = r[0] * U[0] + r[1] * U[1] + r[2] * U[2];
  const double nuv = FAR ? t[12] : n[0] * U[0] + n[1] * U[1] + n[2] * U[2];
  const double fU = k.cU * inv_r, fT = k.cT * inv_r2, fD = k.cD * inv_r2, fS = k.cS * (inv_r2 * inv_r);
  const double rnu = rdn * rdu;
  // displacement
  const double a_t = fU * k.c4;
  const double a_r = fU * rdt - fT * (3.0 * rnu + k.sc6 * nuv);
  const double a_u = -fT * rdn * k.c6;
  const double a_n = fT * k.sc6 * rdu;
#pragma unroll
  for (int i = 0; i < 3; ++i) u[i] = fma(a_n, n[i], fma(a_u, U[i], fma(a_r, r[i], fma(a_t, T[i], u[i]))));
  // stress
  const double k_t = fD * k.c6;
  const double k_rr2 = 1.5 * fD * rdt + fS * (7.5 * rnu - 1.5 * k.c6 * nuv);
  const double k_u = -3.0 * fS * k.nu * rdn;
  const double k_n = -3.0 * fS * k.nu * rdu;
  const double k_nu = -fS * k.c6;
  sdiag += -k_t * rdt - fS * (3.0 * k.c6 * rnu - k.pr4 * nuv);
  double C[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) C[i] = k_t * T[i] + k_u * U[i] + k_n * n[i] + k_rr2 * r[i];
  if (FAR) {
#pragma unroll
    for (int i = 0; i < 3; ++i) s[i] = fma(k_nu, t[13 + i], fma(2.0 * C[i], r[i], s[i]));
    // off-diagonal xy, yz, zx
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int i = q, j = (q + 1) % 3;
      s[3 + q] = fma(k_nu, t[16 + q], fma(C[j], r[i], fma(C[i], r[j], s[3 + q])));
    }
  } else {
    double B[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) B[i] = k_nu * n[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) s[i] = fma(2.0 * B[i], U[i], fma(2.0 * C[i], r[i], s[i]));
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int i = q, j = (q + 1) % 3;
      s[3 + q] = fma(B[j], U[i], fma(B[i], U[j], fma(C[j], r[i], fma(C[i], r[j], s[3 + q]))));
    }
  }
}

// Far pass: grid (point tiles of 64, element slices).
__global__ __launch_bounds__(64) void k_somi_far(int l, int ne, int slice, const double* __restrict__ pts,
                                                  const double* __restrict__ cen,
                                                  const int* __restrict__ far_start,
                                                  const double* __restrict__ tab, SomiConst k,
                                                  double* __restrict__ part) {
  const int ip = blockIdx.x * 64 + threadIdx.x;
  const bool valid = ip < l;
  const int ipr = valid ? ip : l - 1;
  const double px = pts[3 * ipr], py = pts[3 * ipr + 1], pz = pts[3 * ipr + 2];
  const int e0 = blockIdx.y * slice;
  const int e1 = min(ne, e0 + slice);
  double u[3] = {0.0, 0.0, 0.0}, s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, sdiag = 0.0;
  for (int e = e0; e < e1; ++e) {
    const double* c = cen + 4 * e;
    const double dx = px - c[0], dy = py - c[1], dz = pz - c[2];
    if (dx * dx + dy * dy + dz * dz > c[3]) {
      const int g0 = far_start[e], g1 = far_start[e + 1];
      if (g1 > g0) {
        // each Gauss point's row loaded one point ahead, in flight during the
        // previous point's arithmetic
        double cur[kFarStride];
#pragma unroll
        for (int q = 0; q < kFarStride; ++q) cur[q] = tab[(size_t)g0 * kFarStride + q];
#pragma unroll 1
        for (int g = g0; g < g1; ++g) {
          double nxt[kFarStride];
          const double* tq = tab + (size_t)(g + 1 < g1 ? g + 1 : g) * kFarStride;
#pragma unroll
          for (int q = 0; q < kFarStride; ++q) nxt[q] = tq[q];
          somi_gp<true>(cur, px, py, pz, k, u, s, sdiag);
#pragma unroll
          for (int q = 0; q < kFarStride; ++q) cur[q] = nxt[q];
        }
      }
    }
  }
  for (int d = 0; d < 3; ++d) s[d] += sdiag;
  if (valid) {
    double* o = part + ((size_t)blockIdx.y * l + ip) * 9;
    for (int d = 0; d < 3; ++d) o[d] = u[d];
    for (int d = 0; d < 6; ++d) o[3 + d] = s[d];
  }
}

// Near pairs.  A point and an element are near when the point is within the
// element's near threshold of its centroid (stress_mod.f90:584-604, 654).
// The pairs are listed point by point, elements ascending (one wave per
// point, a ballot per 64 elements): k_near_pairs counts them (fill = 0),
// then writes them at the point's offset (fill = 1) with the point of each
// pair in pp.
__global__ __launch_bounds__(64) void k_near_pairs(int l, int ne, const double* __restrict__ pts,
                                                    const double* __restrict__ cen, int fill,
                                                    const int* __restrict__ ptr, int* __restrict__ cnt,
                                                    int* __restrict__ pe, int* __restrict__ pp) {
  const int ip = blockIdx.x;
  if (ip >= l) return;
  const int lane = threadIdx.x;
  const double px = pts[3 * ip], py = pts[3 * ip + 1], pz = pts[3 * ip + 2];
  int run = fill ? ptr[ip] : 0;
  for (int base = 0; base < ne; base += 64) {
    const int e = base + lane;
    bool nearp = false;
    if (e < ne) {
      const double* c = cen + 4 * e;
      const double dx = px - c[0], dy = py - c[1], dz = pz - c[2];
      nearp = !(dx * dx + dy * dy + dz * dz > c[3]);
    }
    const unsigned long long mask = __ballot(nearp);
    if (fill && nearp) {
      const int pos = run + __popcll(mask & ((1ull << lane) - 1ull));
      pe[pos] = e;
      pp[pos] = ip;
    }
    run += __popcll(mask);
  }
  if (!fill && lane == 0) cnt[ip] = run;
}

// Exclusive prefix sum of in[0, n) into out[0, n], the total into out[n]
// and *total: one workgroup, each thread summing a contiguous run.
__global__ __launch_bounds__(1024) void k_scan_excl(int n, const int* __restrict__ in, int* __restrict__ out,
                                                    long long* __restrict__ total) {
  __shared__ long long part[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int b = min(n, t * per), e = min(n, b + per);
  long long sum = 0;
  for (int i = b; i < e; ++i) sum += in[i];
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const long long v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  long long run = part[t] - sum;
  for (int i = b; i < e; ++i) {
    out[i] = (int)run;
    run += in[i];
  }
  if (t == 1023) {
    out[n] = (int)part[1023];
    if (total) *total = part[1023];
  }
}

__global__ void k_zero_int(int n, int* __restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Pairs per element (the order of the atomics does not matter: counts).
__global__ void k_near_ecount(const int* __restrict__ ptr, int l, const int* __restrict__ pe,
                              int* __restrict__ ecnt) {
  const int total = ptr[l];
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += gridDim.x * blockDim.x)
    atomicAdd(ecnt + pe[q], 1);
}

// Per element: its 64-pair work items, and the evaluation counters (far
// points skipped, near points evaluated).
__global__ void k_near_items_count(int ne, const int* __restrict__ nconec, const int* __restrict__ ecnt,
                                   int* __restrict__ cch, unsigned long long* __restrict__ counts) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int c = ecnt[e];
  cch[e] = (c + 63) / 64;
  if (c) {
    atomicAdd(counts + 0, (unsigned long long)c * (nconec[e] == 9 ? 36 : 16));
    atomicAdd(counts + 1, (unsigned long long)c * kNear);
  }
}

// Element-major index of the pairs: ep[eptr[e] ..) holds the pairs of
// element e in an arbitrary order (each pair has its own output row).
__global__ void k_near_efill(const int* __restrict__ ptr, int l, const int* __restrict__ pe,
                             const int* __restrict__ eptr, int* __restrict__ efill, int* __restrict__ ep) {
  const int total = ptr[l];
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += gridDim.x * blockDim.x) {
    const int e = pe[q];
    ep[eptr[e] + atomicAdd(efill + e, 1)] = q;
  }
}

__global__ void k_near_items(int ne, const int* __restrict__ cch, const int* __restrict__ cptr,
                             int2* __restrict__ items) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  for (int c = 0; c < cch[e]; ++c) items[cptr[e] + c] = make_int2(e, 64 * c);
}

// Near pass, element-major: a work item is an element and up to 64 of its
// near points, one per lane; the element's 40x40 table rows are
// wave-uniform (scalar loads), so each table is read once per item.  Each
// lane sums its pair over the table in order and writes the pair's row.
__global__ __launch_bounds__(64) void k_somi_near(const int* __restrict__ cptr, int ne,
                                                   const int2* __restrict__ items,
                                                   const int* __restrict__ eptr, const int* __restrict__ ep,
                                                   const int* __restrict__ pp, const double* __restrict__ pts,
                                                   const double* __restrict__ tab, SomiConst k,
                                                   double* __restrict__ out) {
  const int nitems = cptr[ne];
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const int2 item = items[it];
    const int e = item.x;
    const int j = item.y + (int)threadIdx.x;
    const bool active = eptr[e] + j < eptr[e + 1];
    const int q = active ? ep[eptr[e] + j] : 0;
    const int ip = active ? pp[q] : 0;
    const double px = pts[3 * ip], py = pts[3 * ip + 1], pz = pts[3 * ip + 2];
    double u[3] = {0.0, 0.0, 0.0}, s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, sdiag = 0.0;
    const double* t = tab + (size_t)e * kNear * kSomiStride;
    // each row loaded one point ahead, in flight during the previous
    // point's arithmetic
    double cur[kSomiStride];
#pragma unroll
    for (int qq = 0; qq < kSomiStride; ++qq) cur[qq] = t[qq];
#pragma unroll 1
    for (int g = 0; g < kNear; ++g) {
      double nxt[kSomiStride];
      const double* tq = t + (size_t)(g + 1 < kNear ? g + 1 : g) * kSomiStride;
#pragma unroll
      for (int qq = 0; qq < kSomiStride; ++qq) nxt[qq] = tq[qq];
      somi_gp<false>(cur, px, py, pz, k, u, s, sdiag);
#pragma unroll
      for (int qq = 0; qq < kSomiStride; ++qq) cur[qq] = nxt[qq];
    }
    if (active) {
      double* o = out + (size_t)q * 9;
      for (int d = 0; d < 3; ++d) o[d] = u[d];
      for (int d = 0; d < 3; ++d) o[3 + d] = s[d] + sdiag;
      for (int d = 3; d < 6; ++d) o[3 + d] = s[d];
    }
  }
}

// calc_principal_and_vm, stress_mod.f90:299-359
__device__ void principal_vm(const double* sig, double* prin, double* vm) {
  const double tol = 1.0e-20, pi = 3.14159265358979323846;
  const double sxx = sig[0], syy = sig[1], szz = sig[2], sxy = sig[3], syz = sig[4], szx = sig[5];
  *vm = sqrt(0.5 * ((sxx - syy) * (sxx - syy) + (syy - szz) * (syy - szz) + (szz - sxx) * (szz - sxx) +
                    6.0 * (sxy * sxy + syz * syz + szx * szx)));
  const double i1 = sxx + syy + szz;
  const double q = i1 / 3.0;
  const double pxx = sxx - q, pyy = syy - q, pzz = szz - q;
  const double j2 = 0.5 * (pxx * pxx + pyy * pyy + pzz * pzz) + sxy * sxy + syz * syz + szx * szx;
  if (j2 <= tol * (i1 * i1 + 1.0)) {
    prin[0] = prin[1] = prin[2] = q;
    return;
  }
  const double j3 = pxx * (pyy * pzz - syz * syz) - sxy * (sxy * pzz - syz * szx) + szx * (sxy * syz - pyy * szx);
  const double r = sqrt(4.0 * j2 / 3.0);
  double c3p = 4.0 * j3 / (r * r * r);
  c3p = fmin(1.0, fmax(-1.0, c3p));
  const double phi = acos(c3p) / 3.0;
  double p0 = q + r * cos(phi);
  double p2 = q + r * cos(phi + 2.0 * pi / 3.0);
  double p1 = 3.0 * q - p0 - p2;
  double tmp;
  if (p0 < p1) { tmp = p0; p0 = p1; p1 = tmp; }
  if (p1 < p2) { tmp = p1; p1 = p2; p2 = tmp; }
  if (p0 < p1) { tmp = p0; p0 = p1; p1 = tmp; }
  prin[0] = p0;
  prin[1] = p1;
  prin[2] = p2;
}

__global__ void k_somi_final(int l, int nslice, const double* __restrict__ part,
                             const int* __restrict__ nptr, const double* __restrict__ nearp, double* __restrict__ u_out,
                             double* __restrict__ s_out, double* __restrict__ p_out,
                             double* __restrict__ vm_out) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= l) return;
  double r[9];
  for (int d = 0; d < 9; ++d) r[d] = 0.0;
  for (int sl = 0; sl < nslice; ++sl) {
    const double* o = part + ((size_t)sl * l + ip) * 9;
    for (int d = 0; d < 9; ++d) r[d] += o[d];
  }
  for (int q = nptr[ip]; q < nptr[ip + 1]; ++q)
    for (int d = 0; d < 9; ++d) r[d] += nearp[(size_t)q * 9 + d];
  if (u_out)
    for (int d = 0; d < 3; ++d) u_out[3 * (size_t)ip + d] = r[d];
  if (s_out)
    for (int d = 0; d < 6; ++d) s_out[6 * (size_t)ip + d] = r[3 + d];
  double pr[3], vm;
  principal_vm(r + 3, pr, &vm);
  if (p_out)
    for (int d = 0; d < 3; ++d) p_out[3 * (size_t)ip + d] = pr[d];
  if (vm_out) vm_out[ip] = vm;
}

int somigliana_centres(qd_context* h) {
  QD_ALLOC(h, h->el_centre, 4 * (size_t)h->ne);
  hipLaunchKernelGGL(k_centres, dim3((h->ne + 127) / 128), dim3(128), 0, h->stream, h->ne, h->nconec.p,
                     h->kconec0.p, h->x.p, h->y.p, h->z.p, h->el_centre.p);
  QD_HIP(h, hipGetLastError());
  return QD_OK;
}

int somigliana_run(qd_context* h, int l, const double* pts, const double* fi, const double* dfi,
                   int variant, double* u, double* sig, double* prin, double* vm, long long* counts) {
  const int ne = h->ne;
  {
    double xg[kNearGL], wg[kNearGL];
    gauss_legendre(kNearGL, xg, wg);
    QD_HIP(h, hipStreamSynchronize(h->stream));
    QD_HIP(h, hipMemcpyToSymbol(HIP_SYMBOL(c_gl40x), xg, sizeof(xg)));
    QD_HIP(h, hipMemcpyToSymbol(HIP_SYMBOL(c_gl40w), wg, sizeof(wg)));
  }
  std::vector<int> fs(ne + 1);
  int tot = 0;
  for (int e = 0; e < ne; ++e) {
    fs[e] = tot;
    tot += h->hnconec[e] == 9 ? 36 : 16;
  }
  fs[ne] = tot;
  QD_ALLOC(h, h->far_start, (size_t)ne + 1);
  QD_HIP(h, hipMemcpy(h->far_start.p, fs.data(), fs.size() * sizeof(int), hipMemcpyHostToDevice));
  QD_ALLOC(h, h->far_tab, (size_t)tot * kFarStride);
  QD_ALLOC(h, h->near_tab, (size_t)ne * kNear * kSomiStride);
  QD_ALLOC(h, h->el_centre, 4 * (size_t)ne);
  const int slice = 256;
  const int nslice = (ne + slice - 1) / slice;
  QD_ALLOC(h, h->somi_part, (size_t)nslice * l * 9);
  QD_ALLOC(h, h->counters, 2);
  QD_HIP(h, hipMemsetAsync(h->counters.p, 0, 2 * sizeof(long long), h->stream));

  hipLaunchKernelGGL(k_centres, dim3((ne + 127) / 128), dim3(128), 0, h->stream, ne, h->nconec.p,
                     h->kconec0.p, h->x.p, h->y.p, h->z.p, h->el_centre.p);
  const long long nslot = (long long)ne * kGP;
  hipLaunchKernelGGL(k_far_table, dim3((unsigned)((nslot + 255) / 256)), dim3(256), 0, h->stream, ne,
                     h->nconec.p, h->kconec0.p, h->far_start.p, h->geo.p, fi, dfi, h->far_tab.p);
  const long long nnear = (long long)ne * kNear;
  hipLaunchKernelGGL(k_near_table, dim3((unsigned)((nnear + 255) / 256)), dim3(256), 0, h->stream, ne,
                     h->nconec.p, h->kconec0.p, h->x.p, h->y.p, h->z.p, fi, dfi, h->near_tab.p);
  QD_HIP(h, hipGetLastError());

  const double pi = 3.14159265358979323846;
  const double nu = h->mat.nu, mu = h->mat.mu;
  SomiConst k{};
  k.nu = nu;
  k.c4 = 3.0 - 4.0 * nu;
  k.c6 = 1.0 - 2.0 * nu;
  k.pr4 = 1.0 - 4.0 * nu;
  k.cU = 1.0 / (16.0 * pi * (1.0 - nu) * mu);
  k.cT = -1.0 / (8.0 * pi * (1.0 - nu));
  k.cD = 1.0 / (8.0 * pi * (1.0 - nu));
  k.cS = mu / (4.0 * pi * (1.0 - nu));
  k.sc6 = (variant ? -1.0 : 1.0) * k.c6;
  double* part = h->somi_part.p;
  // FP64 flops of one evaluation as somi_gp computes it (the far form reads
  // n.U and n x U from its table); the work of both passes is set below
  // from the evaluation counters
  constexpr double kFarFlops = 152.0, kNearFlops = 169.0;
  {
    ProfScope ps(h, QD_PROF_SOMIGLIANA_FAR, kFarFlops * (double)tot * l);
    hipLaunchKernelGGL(k_somi_far, dim3((l + 63) / 64, nslice), dim3(64), 0, h->stream, l, ne, slice, pts,
                       h->el_centre.p, h->far_start.p, h->far_tab.p, k, part);
  }
  // near pairs: point-major list (its size sets the allocations), then the
  // element-major index and the work items of the near pass
  QD_ALLOC(h, h->near_cnt, (size_t)l);
  QD_ALLOC(h, h->near_ptr, (size_t)l + 1);
  QD_ALLOC(h, h->near_total, 1);
  hipLaunchKernelGGL(k_near_pairs, dim3(l), dim3(64), 0, h->stream, l, ne, pts, h->el_centre.p, 0,
                     (const int*)nullptr, h->near_cnt.p, (int*)nullptr, (int*)nullptr);
  hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, h->stream, l, h->near_cnt.p, h->near_ptr.p,
                     h->near_total.p);
  QD_HIP(h, hipGetLastError());
  long long npairs = 0;
  QD_HIP(h, hipMemcpyAsync(&npairs, h->near_total.p, sizeof(npairs), hipMemcpyDeviceToHost, h->stream));
  QD_HIP(h, hipStreamSynchronize(h->stream));
  if (npairs > 2147483647LL)
    return set_error(h, QD_ERR_ARG, "compute_internal_points: %lld near point-element pairs", npairs);
  const size_t np = (size_t)std::max<long long>(npairs, 1);
  const size_t nitems_max = (size_t)(npairs / 64) + (size_t)ne;
  QD_ALLOC(h, h->near_pe, np);
  QD_ALLOC(h, h->near_pp, np);
  QD_ALLOC(h, h->near_ep, np);
  QD_ALLOC(h, h->near_out, np * 9);
  QD_ALLOC(h, h->near_ecnt, (size_t)ne);
  QD_ALLOC(h, h->near_efill, (size_t)ne);
  QD_ALLOC(h, h->near_eptr, (size_t)ne + 1);
  QD_ALLOC(h, h->near_cch, (size_t)ne);
  QD_ALLOC(h, h->near_cptr, (size_t)ne + 1);
  QD_ALLOC(h, h->near_items, nitems_max);
  if (npairs > 0) {
    hipLaunchKernelGGL(k_near_pairs, dim3(l), dim3(64), 0, h->stream, l, ne, pts, h->el_centre.p, 1,
                       (const int*)h->near_ptr.p, (int*)nullptr, h->near_pe.p, h->near_pp.p);
    const unsigned eb = (unsigned)((ne + 255) / 256);
    hipLaunchKernelGGL(k_zero_int, dim3(eb), dim3(256), 0, h->stream, ne, h->near_ecnt.p);
    hipLaunchKernelGGL(k_zero_int, dim3(eb), dim3(256), 0, h->stream, ne, h->near_efill.p);
    const unsigned qb = (unsigned)std::min<long long>((npairs + 255) / 256, 4096);
    hipLaunchKernelGGL(k_near_ecount, dim3(qb), dim3(256), 0, h->stream, (const int*)h->near_ptr.p, l,
                       (const int*)h->near_pe.p, h->near_ecnt.p);
    hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, h->stream, ne, h->near_ecnt.p, h->near_eptr.p,
                       (long long*)nullptr);
    hipLaunchKernelGGL(k_near_items_count, dim3(eb), dim3(256), 0, h->stream, ne, h->nconec.p, h->near_ecnt.p,
                       h->near_cch.p, (unsigned long long*)h->counters.p);
    hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, h->stream, ne, h->near_cch.p, h->near_cptr.p,
                       (long long*)nullptr);
    hipLaunchKernelGGL(k_near_efill, dim3(qb), dim3(256), 0, h->stream, (const int*)h->near_ptr.p, l,
                       (const int*)h->near_pe.p, (const int*)h->near_eptr.p, h->near_efill.p, h->near_ep.p);
    hipLaunchKernelGGL(k_near_items, dim3(eb), dim3(256), 0, h->stream, ne, (const int*)h->near_cch.p,
                       (const int*)h->near_cptr.p, h->near_items.p);
    QD_HIP(h, hipGetLastError());
    ProfScope ps(h, QD_PROF_SOMIGLIANA_NEAR, 0.0);
    const unsigned grid = (unsigned)std::min<size_t>(nitems_max, 16384);
    hipLaunchKernelGGL(k_somi_near, dim3(grid), dim3(64), 0, h->stream, (const int*)h->near_cptr.p, ne,
                       (const int2*)h->near_items.p, (const int*)h->near_eptr.p, (const int*)h->near_ep.p,
                       (const int*)h->near_pp.p, pts, h->near_tab.p, k, h->near_out.p);
  }
  hipLaunchKernelGGL(k_somi_final, dim3((l + 127) / 128), dim3(128), 0, h->stream, l, nslice, part,
                     (const int*)h->near_ptr.p, h->near_out.p, u, sig, prin, vm);
  QD_HIP(h, hipGetLastError());
  const bool prof = ((h->prof_mask >> QD_PROF_SOMIGLIANA_FAR) & 1) || ((h->prof_mask >> QD_PROF_SOMIGLIANA_NEAR) & 1);
  if (counts || prof) {
    long long c[2];
    QD_HIP(h, hipMemcpyAsync(c, h->counters.p, sizeof(c), hipMemcpyDeviceToHost, h->stream));
    QD_HIP(h, hipStreamSynchronize(h->stream));
    // far evaluations = all far points of all elements for all points, less
    // those of near elements
    const long long nfar = (long long)tot * l - c[0], nnear = c[1];
    if (counts) {
      counts[0] = nfar;
      counts[1] = nnear;
    }
    // the profile records of this call: the evaluations each pass made
    // (known only now)
    bool far_done = false, near_done = false;
    for (auto it = h->prof_pending.rbegin(); it != h->prof_pending.rend() && !(far_done && near_done); ++it) {
      if (it->kind == QD_PROF_SOMIGLIANA_FAR && !far_done) {
        it->work = kFarFlops * (double)nfar;
        far_done = true;
      } else if (it->kind == QD_PROF_SOMIGLIANA_NEAR && !near_done) {
        it->work = kNearFlops * (double)nnear;
        near_done = true;
      }
    }
  }
  return QD_OK;
}

// ---- boundary stresses (stress_mod.f90:119-292, 366-406) -------------------

__global__ void k_bstress_elem(int ne, const int* __restrict__ nconec, const int* __restrict__ kconec0,
                               const double* __restrict__ x, const double* __restrict__ y,
                               const double* __restrict__ z, const double* __restrict__ fi,
                               const double* __restrict__ dfi, double rge, double xnu,
                               double* __restrict__ se_out) {
  const int ie = blockIdx.x * blockDim.x + threadIdx.x;
  if (ie >= ne) return;
  const int nc = nconec[ie];
  double co[27], ue[27], te[27];
  for (int a = 0; a < nc; ++a) {
    const int ik = kconec0[ie * 9 + a];
    co[3 * a] = x[ik];
    co[3 * a + 1] = y[ik];
    co[3 * a + 2] = z[ik];
    for (int d = 0; d < 3; ++d) {
      ue[3 * a + d] = fi[3 * ik + d];
      te[3 * a + d] = dfi[ie * 27 + 3 * a + d];
    }
  }
  const double q1[9] = {-1.0, 0.0, 1.0, 1.0, 1.0, 0.0, -1.0, -1.0, 0.0};
  const double q2[9] = {-1.0, -1.0, -1.0, 0.0, 1.0, 1.0, 1.0, 0.0, 0.0};
  const double t1[6] = {1.0, 0.0, 0.0, 0.5, 0.0, 0.5};
  const double t2[6] = {0.0, 1.0, 0.0, 0.5, 0.5, 0.0};
  const double c_eps = 2.0 * rge / (1.0 - xnu), c_nu = 2.0 * rge * xnu / (1.0 - xnu), c_sig3 = xnu / (1.0 - xnu);
  double* se = se_out + (size_t)ie * 54;
  for (int i = 0; i < 54; ++i) se[i] = 0.0;
  for (int k = 0; k < nc; ++k) {
    const double g1 = (nc == 9) ? q1[k] : t1[k], g2 = (nc == 9) ? q2[k] : t2[k];
    double f[9], fd1[9], fd2[9];
    shape(nc, g1, g2, f, fd1, fd2);
    double a1[3] = {0, 0, 0}, a2[3] = {0, 0, 0};
    for (int a = 0; a < nc; ++a)
      for (int i = 0; i < 3; ++i) {
        a1[i] += fd1[a] * co[3 * a + i];
        a2[i] += fd2[a] * co[3 * a + i];
      }
    const double v1 = sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
    const double v2 = sqrt(a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2]);
    const double g3[3] = {a1[1] * a2[2] - a1[2] * a2[1], a1[2] * a2[0] - a1[0] * a2[2], a1[0] * a2[1] - a1[1] * a2[0]};
    const double jcb = sqrt(g3[0] * g3[0] + g3[1] * g3[1] + g3[2] * g3[2]);
    double L[9];
    for (int i = 0; i < 3; ++i) {
      L[6 + i] = g3[i] / jcb;
      L[i] = a1[i] / v1;
    }
    L[3] = L[7] * L[2] - L[8] * L[1];
    L[4] = L[8] * L[0] - L[6] * L[2];
    L[5] = L[6] * L[1] - L[7] * L[0];
    double cth = (a1[0] * a2[0] + a1[1] * a2[1] + a1[2] * a2[2]) / (v1 * v2);
    cth = fmin(1.0, fmax(-1.0, cth));
    double sth = sqrt(fmax(0.0, 1.0 - cth * cth));
    if (sth < 1.0e-12) sth = 1.0e-12;
    const double CL11 = 1.0 / v1, CL12 = -cth / (v1 * sth), CL21 = 0.0, CL22 = 1.0 / (v2 * sth);
    double u11 = 0, u22 = 0, u12 = 0, u21 = 0;
    for (int a = 0; a < nc; ++a) {
      const double dn1 = CL11 * fd1[a] + CL21 * fd2[a];
      const double dn2 = CL12 * fd1[a] + CL22 * fd2[a];
      const double ul0 = L[0] * ue[3 * a] + L[1] * ue[3 * a + 1] + L[2] * ue[3 * a + 2];
      const double ul1 = L[3] * ue[3 * a] + L[4] * ue[3 * a + 1] + L[5] * ue[3 * a + 2];
      u11 += dn1 * ul0;
      u22 += dn2 * ul1;
      u12 += dn2 * ul0;
      u21 += dn1 * ul1;
    }
    const double e11 = u11, e22 = u22, e12 = 0.5 * (u12 + u21);
    double tl[3];
    for (int i = 0; i < 3; ++i) tl[i] = L[3 * i] * te[3 * k] + L[3 * i + 1] * te[3 * k + 1] + L[3 * i + 2] * te[3 * k + 2];
    double S[9];
    S[2] = S[6] = tl[0];
    S[5] = S[7] = tl[1];
    S[8] = tl[2];
    S[0] = c_eps * e11 + c_nu * e22 + c_sig3 * S[8];
    S[4] = c_eps * e22 + c_nu * e11 + c_sig3 * S[8];
    S[1] = S[3] = 2.0 * rge * e12;
    double SL[9], Sg[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double acc = 0.0;
        for (int q = 0; q < 3; ++q) acc += S[i * 3 + q] * L[q * 3 + j];
        SL[i * 3 + j] = acc;
      }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double acc = 0.0;
        for (int q = 0; q < 3; ++q) acc += L[q * 3 + i] * SL[q * 3 + j];
        Sg[i * 3 + j] = acc;
      }
    se[6 * k + 0] = Sg[0];
    se[6 * k + 1] = Sg[4];
    se[6 * k + 2] = Sg[8];
    se[6 * k + 3] = 0.5 * (Sg[1] + Sg[3]);
    se[6 * k + 4] = 0.5 * (Sg[5] + Sg[7]);
    se[6 * k + 5] = 0.5 * (Sg[6] + Sg[2]);
  }
}

// Nodal average over the first min(nesim, 8) elements of the node (in
// ascending order), as stress_mod.f90:385-397 averages over ke(1:8, node).
__global__ void k_bstress_node(int n, const int* __restrict__ ke_ptr, const int* __restrict__ ke,
                               const int* __restrict__ nconec, const int* __restrict__ kconec0,
                               const double* __restrict__ se, double* __restrict__ sn,
                               double* __restrict__ pr, double* __restrict__ vm) {
  const int in = blockIdx.x * blockDim.x + threadIdx.x;
  if (in >= n) return;
  double s[6] = {0, 0, 0, 0, 0, 0};
  const int ns = min(ke_ptr[in + 1] - ke_ptr[in], 8);
  for (int j = 0; j < ns; ++j) {
    const int ie = ke[ke_ptr[in] + j] - 1;
    for (int k = 0; k < nconec[ie]; ++k)
      if (kconec0[ie * 9 + k] == in) {
        for (int d = 0; d < 6; ++d) s[d] += se[(size_t)ie * 54 + 6 * k + d];
        break;
      }
  }
  if (ns > 0)
    for (int d = 0; d < 6; ++d) s[d] /= (double)ns;
  for (int d = 0; d < 6; ++d) sn[6 * (size_t)in + d] = s[d];
  principal_vm(s, pr + 3 * (size_t)in, vm + in);
}

int boundary_stress_run(qd_context* h, const double* fi, const double* dfi, double* se, double* sn,
                        double* pr, double* vm) {
  DevBuf<int> dptr, dke;
  QD_ALLOC(h, dptr, (size_t)h->n + 1);
  QD_ALLOC(h, dke, h->ke.size());
  QD_HIP(h, hipMemcpy(dptr.p, h->ke_ptr.data(), (h->n + 1) * sizeof(int), hipMemcpyHostToDevice));
  QD_HIP(h, hipMemcpy(dke.p, h->ke.data(), h->ke.size() * sizeof(int), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_bstress_elem, dim3((h->ne + 63) / 64), dim3(64), 0, h->stream, h->ne, h->nconec.p,
                     h->kconec0.p, h->x.p, h->y.p, h->z.p, fi, dfi, h->mat.mu, h->mat.nu, se);
  hipLaunchKernelGGL(k_bstress_node, dim3((h->n + 127) / 128), dim3(128), 0, h->stream, h->n, dptr.p,
                     dke.p, h->nconec.p, h->kconec0.p, se, sn, pr, vm);
  QD_HIP(h, hipGetLastError());
  QD_HIP(h, hipStreamSynchronize(h->stream));
  return QD_OK;
}

}  // namespace qd
