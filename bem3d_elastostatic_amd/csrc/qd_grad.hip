// qd_grad.hip -- derivatives of the hot path with respect to the node
// coordinates and Poisson's ratio, for the adjoint (reverse-mode) gradient of
// an objective of the internal-point results.  The reference has no
// derivatives; each function differentiated here is cited.
//
//  * ghmat8_vjp_run: d/d(x, y, z, nu) of R = lambda^T (A X - FI) with X held
//    fixed, for ghmat8's A and FI (assembly_mod.f90:53-223).  Let ue and te be
//    the element-node displacements and tractions that X and the boundary
//    values imply (io_mod.f90:148-183: te = DFI; ue = the prescribed value
//    where kode = 0, else the reorganised nodal displacement).  Then R is a sum
//    over (collocation node i, element e) pairs and Gauss points g of
//        w_g [ lambda_i . T*(r; m) (u_h(g) - u_i sum_k N_k) - |m| lambda_i . U*(r) t_h(g) ]
//    with r = y_g - x_i, m = a1 x a2 the unnormalised normal
//    (a_j = sum_k dN_k/dxi_j x_k), u_h, t_h the interpolants of ue, te, and
//    u_i the displacement at the collocation node (ue of the node itself for
//    the singular element).  The u_i term is the rigid-body diagonal
//    (assembly_mod.f90:126-133, 208-217).  T* is linear in the normal, so
//    |m| T*(r; m/|m|) = T*(r; m).  The derivatives in r, m and nu are closed
//    forms (kelvin_pair_grad); the chain to the nodes uses y = sum N_k x_k.
//    The Gauss-point geometry is recomputed from the coordinates.
//  * somigliana_vjp_run: d/d(x, y, z, nu, FI, DFI) of
//        sum_p gu_p . u_p + gs_p . sigma_p
//    for compute_internal_points (stress_mod.f90:511-781): dual numbers over
//    the Gauss-point integrand of accumulate_somigliana_range
//    (stress_mod.f90:807-885), seeded on (r, m, nu) and on the weighted
//    boundary interpolants (u_h w, t_h w).  Each work item is one Gauss point
//    of one element; it sweeps the internal points the forward integrates it
//    against (far points with the element rule, near points with the 40x40
//    rule, split by the same centroid test, stress_mod.f90:584-604).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "qd_device.h"
#include "qd_internal.h"

namespace qd {
namespace {

constexpr double kPi = 3.14159265358979323846;

// assembly_mod.f90:6-19: the rules the forward pass integrates with.
__constant__ double d_gi[6] = {0.238619186083197, -0.238619186083197, 0.661209386466265,
                               -0.661209386466265, 0.932469514203152, -0.932469514203152};
__constant__ double d_ome[6] = {0.467913934572691, 0.467913934572691, 0.360761573048139,
                                0.360761573048139, 0.171324492379170, 0.171324492379170};
__constant__ double d_rr[4] = {0.0694318422, 0.3300094782, 0.6699905218, 0.9305681558};
__constant__ double d_ss[4] = {0.0571041961, 0.2768430136, 0.5835904324, 0.8602401357};
__constant__ double d_omei[4] = {0.1739274226, 0.3260725774, 0.3260725774, 0.1739274226};
__constant__ double d_omej[4] = {0.1355069134, 0.2034645680, 0.1298475476, 0.0311809709};
__constant__ double d_gl40x[kNearGL];
__constant__ double d_gl40w[kNearGL];

constexpr int kNearPts = kNearGL * kNearGL;  // 1600 = 25 waves
constexpr int kAcc = 13;                     // dr(3) dm(3) dnu du_w(3) dt_w(3)

// ---- forward-mode dual numbers ---------------------------------------------
template <int K>
struct Dual {
  double v;
  double d[K];
};

template <int K>
__device__ __forceinline__ Dual<K> dconst(double v) {
  Dual<K> r;
  r.v = v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = 0.0;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> dseed(double v, int slot) {
  Dual<K> r = dconst<K>(v);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = (k == slot) ? 1.0 : 0.0;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator+(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a) {
  Dual<K> r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator+(const Dual<K>& a, double b) {
  Dual<K> r = a;
  r.v += b;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator+(double a, const Dual<K>& b) {
  return b + a;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a, double b) {
  Dual<K> r = a;
  r.v -= b;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(double a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a - b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = -b.d[k];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(const Dual<K>& a, double b) {
  Dual<K> r;
  r.v = a.v * b;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(double a, const Dual<K>& b) {
  return b * a;
}
template <int K>
__device__ __forceinline__ Dual<K> recip(const Dual<K>& a) {
  const double iv = 1.0 / a.v;
  Dual<K> r;
  r.v = iv;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = -a.d[k] * iv * iv;
  return r;
}
__device__ __forceinline__ double recip(double a) { return 1.0 / a; }
template <int K>
__device__ __forceinline__ Dual<K> operator/(const Dual<K>& a, const Dual<K>& b) {
  return a * recip(b);
}
template <int K>
__device__ __forceinline__ Dual<K> operator/(const Dual<K>& a, double b) {
  return a * (1.0 / b);
}
template <int K>
__device__ __forceinline__ Dual<K> operator/(double a, const Dual<K>& b) {
  return a * recip(b);
}
template <int K>
__device__ __forceinline__ Dual<K> dsqrt(const Dual<K>& a) {
  const double s = sqrt(a.v);
  const double h = 0.5 / s;
  Dual<K> r;
  r.v = s;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * h;
  return r;
}
__device__ __forceinline__ double dsqrt(double a) { return sqrt(a); }

// The Gauss-point term of accumulate_somigliana_range (stress_mod.f90:807-885)
// contracted with (gu, gs).  r = y_g - x_p; m = the unnormalised normal;
// uw, tw = the boundary displacement and traction interpolants times the
// quadrature weight without the Jacobian (the Jacobian is |m|); sg = +1 keeps
// the reference's sign of the displacement T* term, -1 the textbook one.
// TG and TV are double or dual numbers; each seeds its own slots.
template <typename TG, typename TV>
__device__ __forceinline__ auto somi_term(const TG (&r)[3], const TG (&m)[3], const TG& nu, const TV (&uw)[3],
                                          const TV (&tw)[3], double mu, double sg, const double* __restrict__ gu,
                                          const double* __restrict__ gs) {
  const auto omn = 1.0 - nu;
  const auto c4 = 3.0 - 4.0 * nu;
  const auto c6 = 1.0 - 2.0 * nu;
  const auto pr4 = 1.0 - 4.0 * nu;
  const auto iomn = recip(omn);
  const auto cU = (1.0 / (16.0 * kPi * mu)) * iomn;
  const auto cT = (-1.0 / (8.0 * kPi)) * iomn;
  const auto cD = (1.0 / (8.0 * kPi)) * iomn;
  const auto cS = (mu / (4.0 * kPi)) * iomn;
  const auto jac = dsqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  const auto ij = recip(jac);
  const auto n1 = m[0] * ij, n2 = m[1] * ij, n3 = m[2] * ij;
  const auto tu1 = uw[0] * jac, tu2 = uw[1] * jac, tu3 = uw[2] * jac;
  const auto tt1 = tw[0] * jac, tt2 = tw[1] * jac, tt3 = tw[2] * jac;
  const auto inv_r = recip(dsqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]));
  const auto inv_r2 = inv_r * inv_r;
  const auto inv_r3 = inv_r2 * inv_r;
  const auto rd1 = r[0] * inv_r, rd2 = r[1] * inv_r, rd3 = r[2] * inv_r;
  const auto rdn = rd1 * n1 + rd2 * n2 + rd3 * n3;
  const auto rdt = rd1 * tt1 + rd2 * tt2 + rd3 * tt3;
  const auto rdu = rd1 * tu1 + rd2 * tu2 + rd3 * tu3;
  const auto nuv = n1 * tu1 + n2 * tu2 + n3 * tu3;
  const auto fU = cU * inv_r;
  const auto fT = cT * inv_r2;
  const auto sc6 = sg * c6;
  const auto u0 = fU * (c4 * tt1 + rd1 * rdt) - fT * (rdn * (c6 * tu1 + 3.0 * rd1 * rdu) + sc6 * (nuv * rd1 - n1 * rdu));
  const auto u1 = fU * (c4 * tt2 + rd2 * rdt) - fT * (rdn * (c6 * tu2 + 3.0 * rd2 * rdu) + sc6 * (nuv * rd2 - n2 * rdu));
  const auto u2 = fU * (c4 * tt3 + rd3 * rdt) - fT * (rdn * (c6 * tu3 + 3.0 * rd3 * rdu) + sc6 * (nuv * rd3 - n3 * rdu));
  const auto fD = cD * inv_r2;
  const auto fS = cS * inv_r3;
  const auto Dxx = fD * (c6 * (2.0 * tt1 * rd1 - rdt) + 3.0 * rd1 * rd1 * rdt);
  const auto Dyy = fD * (c6 * (2.0 * tt2 * rd2 - rdt) + 3.0 * rd2 * rd2 * rdt);
  const auto Dzz = fD * (c6 * (2.0 * tt3 * rd3 - rdt) + 3.0 * rd3 * rd3 * rdt);
  const auto Dxy = fD * (c6 * (tt1 * rd2 + tt2 * rd1) + 3.0 * rd1 * rd2 * rdt);
  const auto Dyz = fD * (c6 * (tt2 * rd3 + tt3 * rd2) + 3.0 * rd2 * rd3 * rdt);
  const auto Dzx = fD * (c6 * (tt3 * rd1 + tt1 * rd3) + 3.0 * rd3 * rd1 * rdt);
  const auto Sxx = fS * (3.0 * rdn * (c6 * rdu + 2.0 * nu * tu1 * rd1 - 5.0 * rd1 * rd1 * rdu) +
                         6.0 * nu * rdu * n1 * rd1 + c6 * (3.0 * nuv * rd1 * rd1 + 2.0 * n1 * tu1) - pr4 * nuv);
  const auto Syy = fS * (3.0 * rdn * (c6 * rdu + 2.0 * nu * tu2 * rd2 - 5.0 * rd2 * rd2 * rdu) +
                         6.0 * nu * rdu * n2 * rd2 + c6 * (3.0 * nuv * rd2 * rd2 + 2.0 * n2 * tu2) - pr4 * nuv);
  const auto Szz = fS * (3.0 * rdn * (c6 * rdu + 2.0 * nu * tu3 * rd3 - 5.0 * rd3 * rd3 * rdu) +
                         6.0 * nu * rdu * n3 * rd3 + c6 * (3.0 * nuv * rd3 * rd3 + 2.0 * n3 * tu3) - pr4 * nuv);
  const auto Sxy = fS * (3.0 * rdn * (nu * (tu1 * rd2 + tu2 * rd1) - 5.0 * rd1 * rd2 * rdu) +
                         3.0 * nu * rdu * (n1 * rd2 + n2 * rd1) + c6 * (3.0 * nuv * rd1 * rd2 + n1 * tu2 + n2 * tu1));
  const auto Syz = fS * (3.0 * rdn * (nu * (tu2 * rd3 + tu3 * rd2) - 5.0 * rd2 * rd3 * rdu) +
                         3.0 * nu * rdu * (n2 * rd3 + n3 * rd2) + c6 * (3.0 * nuv * rd2 * rd3 + n2 * tu3 + n3 * tu2));
  const auto Szx = fS * (3.0 * rdn * (nu * (tu3 * rd1 + tu1 * rd3) - 5.0 * rd3 * rd1 * rdu) +
                         3.0 * nu * rdu * (n3 * rd1 + n1 * rd3) + c6 * (3.0 * nuv * rd3 * rd1 + n3 * tu1 + n1 * tu3));
  return gu[0] * u0 + gu[1] * u1 + gu[2] * u2 + gs[0] * (Dxx - Sxx) + gs[1] * (Dyy - Syy) + gs[2] * (Dzz - Szz) +
         gs[3] * (Dxy - Sxy) + gs[4] * (Dyz - Syz) + gs[5] * (Dzx - Szx);
}

// ---- Gauss-point geometry, recomputed from the coordinates ------------------
struct GpGeom {
  double f[9], fd1[9], fd2[9];
  double y[3], a1[3], a2[3];
};

// Element rule point k (etaxja's points and weights, assembly_mod.f90:225-337).
__device__ __forceinline__ void element_rule(int nc, int k, double& g1, double& g2, double& wq) {
  if (nc == 9) {
    const int i = k / 6, j = k % 6;
    g1 = d_gi[i];
    g2 = d_gi[j];
    wq = d_ome[i] * d_ome[j];
  } else {
    const int i = k / 4, j = k % 4;
    g1 = d_ss[j];
    g2 = d_rr[i] * (1.0 - d_ss[j]);
    wq = d_omei[i] * d_omej[j];
  }
}

// 40x40 near rule point g (stress_mod.f90:659-680); wq excludes the Jacobian.
__device__ __forceinline__ void near_rule(int nc, int g, double& g1, double& g2, double& wq) {
  const int i = g / kNearGL, j = g % kNearGL;
  if (nc == 9) {
    g1 = d_gl40x[i];
    g2 = d_gl40x[j];
    wq = d_gl40w[i] * d_gl40w[j];
  } else {
    g1 = 0.5 * (1.0 + d_gl40x[i]);
    g2 = 0.25 * (1.0 - d_gl40x[i]) * (1.0 + d_gl40x[j]);
    wq = d_gl40w[i] * d_gl40w[j] * 0.125 * (1.0 - d_gl40x[i]);
  }
}

__device__ __forceinline__ void gp_geometry(int nc, double g1, double g2, const int* __restrict__ nodes,
                                            const double* __restrict__ x, const double* __restrict__ y,
                                            const double* __restrict__ z, GpGeom& G) {
  shape(nc, g1, g2, G.f, G.fd1, G.fd2);
#pragma unroll
  for (int d = 0; d < 3; ++d) G.y[d] = G.a1[d] = G.a2[d] = 0.0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    if (k < nc) {
      const int nd = nodes[k];
      const double xk[3] = {x[nd], y[nd], z[nd]};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        G.y[d] += G.f[k] * xk[d];
        G.a1[d] += G.fd1[k] * xk[d];
        G.a2[d] += G.fd2[k] * xk[d];
      }
    }
  }
}

__device__ __forceinline__ void cross(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Contribution of one Gauss point with weight derivatives (gr = dS/dy,
// gm = dS/dm) to node k of the element: N_k gr + dN_k/dxi1 (a2 x gm) +
// dN_k/dxi2 (gm x a1), since y = sum N_k x_k and m = a1 x a2.
__device__ __forceinline__ void chain_node(const GpGeom& G, int k, const double* gr, const double* c1,
                                           const double* c2, double* out) {
#pragma unroll
  for (int d = 0; d < 3; ++d) out[d] = G.f[k] * gr[d] + G.fd1[k] * c1[d] + G.fd2[k] * c2[d];
}

// ---- assembly residual ---------------------------------------------------------
// d/d(r, m, nu) of  wq [ lam . T*(r; m) dl - |m| lam . U*(r) t ]  with the
// Kelvin kernels of funkelvin (assembly_mod.f90:618-650):
//   lam . T*(r; m) dl = c5 [ c6 p q / r^3 + 3 q a b / r^5 + c6 (lm b - dm a) / r^3 ]
//   lam . U*(r) t     = c3 [ c4 f / r + a e / r^3 ]
// with a = lam.r, b = dl.r, q = r.m, p = lam.dl, lm = lam.m, dm = dl.m,
// f = lam.t, e = t.r.
__device__ __forceinline__ void kelvin_pair_grad(const double* r, const double* m, const double* lam,
                                                 const double* dl, const double* t, double wq, const Material& mt,
                                                 double* gr, double* gm, double& gnu) {
  const double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  const double p1 = inv_sqrt(rr);
  const double p2 = p1 * p1, p3 = p2 * p1, p5 = p3 * p2, p7 = p5 * p2;
  const double a = lam[0] * r[0] + lam[1] * r[1] + lam[2] * r[2];
  const double b = dl[0] * r[0] + dl[1] * r[1] + dl[2] * r[2];
  const double q = r[0] * m[0] + r[1] * m[1] + r[2] * m[2];
  const double p = lam[0] * dl[0] + lam[1] * dl[1] + lam[2] * dl[2];
  const double lm = lam[0] * m[0] + lam[1] * m[1] + lam[2] * m[2];
  const double dm = dl[0] * m[0] + dl[1] * m[1] + dl[2] * m[2];
  const double f = lam[0] * t[0] + lam[1] * t[1] + lam[2] * t[2];
  const double e = t[0] * r[0] + t[1] * r[1] + t[2] * r[2];
  const double J = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  const double c3 = mt.c3, c4 = mt.c4, c5 = mt.c5, c6 = mt.c6;
  const double P = c5 * wq, Q = c3 * wq;
  const double anti = lm * b - dm * a;
  const double h0 = c6 * p * q * p3 + 3.0 * q * a * b * p5 + c6 * anti * p3;
  const double g0 = c4 * f * p1 + a * e * p3;
  // H part
  const double hm_r = c6 * p * p3 + 3.0 * a * b * p5;
  const double hr_r = -3.0 * c6 * p * q * p5 - 15.0 * q * a * b * p7 - 3.0 * c6 * anti * p5;
  const double hr_m = c6 * p * p3 + 3.0 * a * b * p5;
  // G part
  const double gr_r = Q * J * (-c4 * f * p3 - 3.0 * a * e * p5);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double dHm = P * (hm_r * r[d] + c6 * p3 * (b * lam[d] - a * dl[d]));
    const double dHr = P * (hr_r * r[d] + hr_m * m[d] + 3.0 * p5 * q * (b * lam[d] + a * dl[d]) +
                            c6 * p3 * (lm * dl[d] - dm * lam[d]));
    const double dGm = Q * g0 * m[d] / J;
    const double dGr = gr_r * r[d] + Q * J * p3 * (e * lam[d] + a * t[d]);
    gm[d] = dHm - dGm;
    gr[d] = dHr - dGr;
  }
  const double iomn = 1.0 / (1.0 - mt.nu);
  const double dH = wq * h0 * (c5 * iomn) + P * (p * q * p3 + anti * p3) * (-2.0);
  const double dG = wq * J * g0 * (c3 * iomn) + Q * J * f * p1 * (-4.0);
  gnu = dH - dG;
}

struct AvjpArgs {
  int n, ne;
  Material mat;
  const double *x, *y, *z;
  const int* nconec;
  const int* kconec0;
  const double* lam;  // (3n)
  const double* uc;   // (3n) reorganised nodal displacement
  const double* ue;   // (ne, 27) element-node displacement
  const double* te;   // (ne, 27) element-node traction (DFI)
};

constexpr int kRec = 20;  // y(3) a1(3) a2(3) uh(3) th(3) sumN wq pad(3)

// Per (element, rule point): geometry and interpolants of ue, te.
__global__ void k_avjp_table(AvjpArgs a, double* __restrict__ rec) {
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= (long long)a.ne * kGP) return;
  const int e = (int)(tid / kGP), k = (int)(tid % kGP);
  const int nc = a.nconec[e];
  double* R = rec + tid * kRec;
  if (k >= (nc == 9 ? 36 : 16)) {
    for (int q = 0; q < kRec; ++q) R[q] = 0.0;
    return;
  }
  double g1, g2, wq;
  element_rule(nc, k, g1, g2, wq);
  GpGeom G;
  gp_geometry(nc, g1, g2, a.kconec0 + e * 9, a.x, a.y, a.z, G);
  double uh[3] = {0, 0, 0}, th[3] = {0, 0, 0}, sn = 0.0;
  for (int m = 0; m < nc; ++m) {
    sn += G.f[m];
    for (int d = 0; d < 3; ++d) {
      uh[d] += G.f[m] * a.ue[e * 27 + 3 * m + d];
      th[d] += G.f[m] * a.te[e * 27 + 3 * m + d];
    }
  }
  for (int d = 0; d < 3; ++d) {
    R[d] = G.y[d];
    R[3 + d] = G.a1[d];
    R[6 + d] = G.a2[d];
    R[9 + d] = uh[d];
    R[12 + d] = th[d];
  }
  R[15] = sn;
  R[16] = wq;
  R[17] = R[18] = R[19] = 0.0;
}

// Shape functions and derivatives at the element rule points: [kind][k][27].
__global__ void k_shape_table(double* __restrict__ shp) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * kGP) return;
  const int kind = t / kGP, k = t % kGP;
  const int nc = kind == 0 ? 9 : 6;
  double* S = shp + (size_t)t * 27;
  if (k >= (nc == 9 ? 36 : 16)) {
    for (int q = 0; q < 27; ++q) S[q] = 0.0;
    return;
  }
  double g1, g2, wq;
  element_rule(nc, k, g1, g2, wq);
  double f[9], fd1[9], fd2[9];
  shape(nc, g1, g2, f, fd1, fd2);
  for (int m = 0; m < 9; ++m) {
    S[m] = f[m];
    S[9 + m] = fd1[m];
    S[18 + m] = fd2[m];
  }
}

// Sum over the wave of acc[27] into the element's nodes (lane q sums row q).
__device__ __forceinline__ void scatter_nodes27(double (&acc)[27], double (*sred)[65], const int* __restrict__ nodes,
                                                int nc, double* __restrict__ gx, double* __restrict__ gy,
                                                double* __restrict__ gz) {
  const int lane = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 27; ++q) sred[q][lane] = acc[q];
  __syncthreads();
  if (lane < 3 * nc) {
    double s = 0.0;
    for (int j = 0; j < 64; ++j) s += sred[lane][j];
    const int nd = nodes[lane / 3];
    double* g = (lane % 3 == 0) ? gx : (lane % 3 == 1) ? gy : gz;
    atomicAdd(g + nd, s);
  }
  __syncthreads();
}

constexpr int kAvjpChunk = 16;

// Regular pairs: one lane per collocation node, a chunk of elements per wave.
__global__ __launch_bounds__(64) void k_avjp_regular(AvjpArgs a, const double* __restrict__ rec,
                                                     const double* __restrict__ shp, double* __restrict__ gx,
                                                     double* __restrict__ gy, double* __restrict__ gz,
                                                     double* __restrict__ gnu) {
  __shared__ double sred[27][65];
  const int lane = threadIdx.x;
  const int i = blockIdx.x * 64 + lane;
  const bool valid = i < a.n;
  const int ic = valid ? i : a.n - 1;
  const double xi[3] = {a.x[ic], a.y[ic], a.z[ic]};
  const double lam[3] = {a.lam[3 * ic], a.lam[3 * ic + 1], a.lam[3 * ic + 2]};
  const double uc[3] = {a.uc[3 * ic], a.uc[3 * ic + 1], a.uc[3 * ic + 2]};
  double accx[3] = {0.0, 0.0, 0.0}, accnu = 0.0;
  const int e0 = blockIdx.y * kAvjpChunk;
  const int e1 = min(a.ne, e0 + kAvjpChunk);
  for (int e = e0; e < e1; ++e) {
    const int nc = a.nconec[e];
    const int kind = nc == 9 ? 0 : 1;
    const int ngp = nc == 9 ? 36 : 16;
    const int* nodes = a.kconec0 + e * 9;
    bool own = false;
    for (int k = 0; k < nc; ++k) own |= nodes[k] == i;
    const bool act = valid && !own;
    double acc[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) acc[q] = 0.0;
    for (int g = 0; g < ngp; ++g) {
      const double* R = rec + ((size_t)e * kGP + g) * kRec;
      const double* S = shp + ((size_t)kind * kGP + g) * 27;
      if (act) {
        const double r[3] = {R[0] - xi[0], R[1] - xi[1], R[2] - xi[2]};
        double m[3];
        cross(R + 3, R + 6, m);
        const double dl[3] = {R[9] - uc[0] * R[15], R[10] - uc[1] * R[15], R[11] - uc[2] * R[15]};
        const double t[3] = {R[12], R[13], R[14]};
        double gr[3], gm[3], gn;
        kelvin_pair_grad(r, m, lam, dl, t, R[16], a.mat, gr, gm, gn);
        accnu += gn;
        double c1[3], c2[3];
        cross(R + 6, gm, c1);
        cross(gm, R + 3, c2);
#pragma unroll
        for (int d = 0; d < 3; ++d) accx[d] -= gr[d];
#pragma unroll
        for (int k = 0; k < 9; ++k)
#pragma unroll
          for (int d = 0; d < 3; ++d) acc[3 * k + d] += S[k] * gr[d] + S[9 + k] * c1[d] + S[18 + k] * c2[d];
      }
    }
    scatter_nodes27(acc, sred, nodes, nc, gx, gy, gz);
  }
  if (valid) {
    atomicAdd(gx + i, accx[0]);
    atomicAdd(gy + i, accx[1]);
    atomicAdd(gz + i, accx[2]);
  }
  const double sn = wave_sum(accnu);
  if (lane == 0) atomicAdd(gnu, sn);
}

// Singular pairs (locin8/triloc8, assembly_mod.f90:386-616): one wave per
// (collocation node, own element); lanes over the sub-triangle rule points.
__global__ __launch_bounds__(64) void k_avjp_singular(AvjpArgs a, int npairs, const SingPair* __restrict__ pairs,
                                                      const SubTri* __restrict__ subtri, double* __restrict__ gx,
                                                      double* __restrict__ gy, double* __restrict__ gz,
                                                      double* __restrict__ gnu) {
  __shared__ double sred[27][65];
  const int p = blockIdx.x;
  if (p >= npairs) return;
  const int lane = threadIdx.x;
  const SingPair sp = pairs[p];
  const int e = sp.elem, nc = a.nconec[e];
  const int kind = nc == 9 ? 0 : 1;
  const int* nodes = a.kconec0 + e * 9;
  const int io = sp.nodo - 1;
  const int in = sp.node;
  const double xi[3] = {a.x[in], a.y[in], a.z[in]};
  const double lam[3] = {a.lam[3 * in], a.lam[3 * in + 1], a.lam[3 * in + 2]};
  const double* ue = a.ue + e * 27;
  const double* te = a.te + e * 27;
  const double uo[3] = {ue[3 * io], ue[3 * io + 1], ue[3 * io + 2]};
  double acc[27];
#pragma unroll
  for (int q = 0; q < 27; ++q) acc[q] = 0.0;
  double accnu = 0.0;
  const int ngp = sp.nsub * 36;
  for (int g = lane; g < ngp; g += 64) {
    const SubTri st = subtri[(kind * 10 + sp.nodo) * 4 + g / 36];
    const int i = (g % 36) / 6, j = g % 6;
    const double ss1 = 0.5 * (d_gi[i] + 1.0), ss2 = 0.5 * (d_gi[j] + 1.0);
    const double g1 = (1.0 - ss1) * st.v1[0] + ss1 * (1.0 - ss2) * st.v1[1] + ss1 * ss2 * st.v1[2];
    const double g2 = (1.0 - ss1) * st.v2[0] + ss1 * (1.0 - ss2) * st.v2[1] + ss1 * ss2 * st.v2[2];
    const double wq = (d_gi[i] + 1.0) * st.area * 0.25 * d_ome[i] * d_ome[j];
    GpGeom G;
    gp_geometry(nc, g1, g2, nodes, a.x, a.y, a.z, G);
    double uh[3] = {0, 0, 0}, th[3] = {0, 0, 0}, sn = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      if (k < nc) {
        sn += G.f[k];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          uh[d] += G.f[k] * ue[3 * k + d];
          th[d] += G.f[k] * te[3 * k + d];
        }
      }
    }
    const double r[3] = {G.y[0] - xi[0], G.y[1] - xi[1], G.y[2] - xi[2]};
    double m[3];
    cross(G.a1, G.a2, m);
    const double dl[3] = {uh[0] - uo[0] * sn, uh[1] - uo[1] * sn, uh[2] - uo[2] * sn};
    double gr[3], gm[3], gn;
    kelvin_pair_grad(r, m, lam, dl, th, wq, a.mat, gr, gm, gn);
    accnu += gn;
    double c1[3], c2[3];
    cross(G.a2, gm, c1);
    cross(gm, G.a1, c2);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      double o[3];
      chain_node(G, k, gr, c1, c2, o);
#pragma unroll
      for (int d = 0; d < 3; ++d) acc[3 * k + d] += o[d] - ((k == io) ? gr[d] : 0.0);
    }
  }
  scatter_nodes27(acc, sred, nodes, nc, gx, gy, gz);
  const double s = wave_sum(accnu);
  if (lane == 0) atomicAdd(gnu, s);
}

// ---- Somigliana ---------------------------------------------------------------------
struct SvjpArgs {
  int l, ne;
  double mu, nu, sg;
  const double* pts;  // (l, 3)
  const double* gu;   // (l, 3)
  const double* gs;   // (l, 6)
  const double* cen;  // (ne, 4): centroid, near threshold
  const int* nconec;
  const int* kconec0;
  const double *x, *y, *z;
  const double* fi;   // (3n)
  const double* dfi;  // (ne, 27)
};

// Interpolants of the boundary displacement and traction at a rule point.
__device__ __forceinline__ void boundary_values(const SvjpArgs& a, int e, int nc, const GpGeom& G, double wq,
                                                double* uw, double* tw) {
  double uh[3] = {0, 0, 0}, th[3] = {0, 0, 0};
  const int* nodes = a.kconec0 + e * 9;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    if (k < nc) {
      const int nd = nodes[k];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        uh[d] += G.f[k] * a.fi[3 * nd + d];
        th[d] += G.f[k] * a.dfi[e * 27 + 3 * k + d];
      }
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    uw[d] = uh[d] * wq;
    tw[d] = th[d] * wq;
  }
}

// Derivatives of one (rule point, internal point) term, one group of slots
// per call (small duals keep the integrand in registers):
//   G = 0: d/dr -> acc[0..2]    G = 1: d/dm -> acc[3..5]    G = 2: d/dnu -> acc[6]
//   G = 3: d/duw -> acc[7..9]   G = 4: d/dtw -> acc[10..12]
template <int G>
__device__ __noinline__ void somi_group(const SvjpArgs& a, int p, const double* yg, const double* m,
                                           const double* uw, const double* tw, double* acc) {
  const double* gu = a.gu + 3 * (size_t)p;
  const double* gs = a.gs + 6 * (size_t)p;
  const double rv[3] = {yg[0] - a.pts[3 * (size_t)p], yg[1] - a.pts[3 * (size_t)p + 1],
                        yg[2] - a.pts[3 * (size_t)p + 2]};
  if constexpr (G <= 2) {
    constexpr int K = (G == 2) ? 1 : 3;
    Dual<K> r[3], md[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      r[d] = (G == 0) ? dseed<K>(rv[d], d) : dconst<K>(rv[d]);
      md[d] = (G == 1) ? dseed<K>(m[d], d) : dconst<K>(m[d]);
    }
    const Dual<K> nu = (G == 2) ? dseed<K>(a.nu, 0) : dconst<K>(a.nu);
    const double uwv[3] = {uw[0], uw[1], uw[2]}, twv[3] = {tw[0], tw[1], tw[2]};
    const Dual<K> s = somi_term(r, md, nu, uwv, twv, a.mu, a.sg, gu, gs);
    constexpr int off = (G == 0) ? 0 : (G == 1) ? 3 : 6;
#pragma unroll
    for (int k = 0; k < K; ++k) acc[off + k] += s.d[k];
  } else {
    const double md[3] = {m[0], m[1], m[2]};
    Dual<3> uwd[3], twd[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      uwd[d] = (G == 3) ? dseed<3>(uw[d], d) : dconst<3>(uw[d]);
      twd[d] = (G == 4) ? dseed<3>(tw[d], d) : dconst<3>(tw[d]);
    }
    const double nu = a.nu;
    const Dual<3> s = somi_term(rv, md, nu, uwd, twd, a.mu, a.sg, gu, gs);
    constexpr int off = (G == 3) ? 7 : 10;
#pragma unroll
    for (int k = 0; k < 3; ++k) acc[off + k] += s.d[k];
  }
}

__device__ __forceinline__ void somi_point(const SvjpArgs& a, int p, const double* yg, const double* m,
                                           const double* uw, const double* tw, double* acc) {
  somi_group<0>(a, p, yg, m, uw, tw, acc);
  somi_group<1>(a, p, yg, m, uw, tw, acc);
  somi_group<2>(a, p, yg, m, uw, tw, acc);
  somi_group<3>(a, p, yg, m, uw, tw, acc);
  somi_group<4>(a, p, yg, m, uw, tw, acc);
}

// Rule point state the point sweeps need: y, m, uw, tw.
__device__ __forceinline__ void rule_point_state(const SvjpArgs& a, int e, int nc, double g1, double g2, double wq,
                                                 double* yg, double* m, double* uw, double* tw) {
  GpGeom G;
  gp_geometry(nc, g1, g2, a.kconec0 + e * 9, a.x, a.y, a.z, G);
  cross(G.a1, G.a2, m);
  boundary_values(a, e, nc, G, wq, uw, tw);
#pragma unroll
  for (int d = 0; d < 3; ++d) yg[d] = G.y[d];
}

// Far terms: one lane per element rule point, a chunk of internal points per
// workgroup column; per-point sums go to acc (nitems, 13) with atomics.
__global__ __launch_bounds__(64) void k_svjp_far(SvjpArgs a, int nitems, const int* __restrict__ item_elem,
                                                 const int* __restrict__ far_start, int pchunk,
                                                 double* __restrict__ accout) {
  const int it = blockIdx.x * 64 + threadIdx.x;
  const bool valid = it < nitems;
  const int itc = valid ? it : nitems - 1;
  const int e = item_elem[itc];
  const int k = itc - far_start[e];
  const int nc = a.nconec[e];
  double g1, g2, wq;
  element_rule(nc, k, g1, g2, wq);
  double yg[3], m[3], uw[3], tw[3];
  rule_point_state(a, e, nc, g1, g2, wq, yg, m, uw, tw);
  const double cx = a.cen[4 * e], cy = a.cen[4 * e + 1], cz = a.cen[4 * e + 2], thr = a.cen[4 * e + 3];
  double acc[kAcc];
#pragma unroll
  for (int q = 0; q < kAcc; ++q) acc[q] = 0.0;
  const int p0 = blockIdx.y * pchunk;
  const int p1 = min(a.l, p0 + pchunk);
  for (int p = p0; p < p1; ++p) {
    const double dx = a.pts[3 * (size_t)p] - cx, dy = a.pts[3 * (size_t)p + 1] - cy,
                 dz = a.pts[3 * (size_t)p + 2] - cz;
    if (valid && dx * dx + dy * dy + dz * dz > thr) somi_point(a, p, yg, m, uw, tw, acc);
  }
  if (valid)
    for (int q = 0; q < kAcc; ++q) atomicAdd(accout + (size_t)it * kAcc + q, acc[q]);
}

// Far chain: per element rule point, to node coordinates, FI and DFI.
__global__ __launch_bounds__(64) void k_svjp_far_chain(SvjpArgs a, int nitems, const int* __restrict__ item_elem,
                                                       const int* __restrict__ far_start,
                                                       const double* __restrict__ accin, double* __restrict__ gx,
                                                       double* __restrict__ gy, double* __restrict__ gz,
                                                       double* __restrict__ gfi, double* __restrict__ gdfi,
                                                       double* __restrict__ gnu) {
  const int it = blockIdx.x * 64 + threadIdx.x;
  double nu_part = 0.0;
  if (it < nitems) {
    const int e = item_elem[it];
    const int k = it - far_start[e];
    const int nc = a.nconec[e];
    double g1, g2, wq;
    element_rule(nc, k, g1, g2, wq);
    GpGeom G;
    const int* nodes = a.kconec0 + e * 9;
    gp_geometry(nc, g1, g2, nodes, a.x, a.y, a.z, G);
    const double* A = accin + (size_t)it * kAcc;
    double c1[3], c2[3];
    cross(G.a2, A + 3, c1);
    cross(A + 3, G.a1, c2);
    nu_part = A[6];
    for (int q = 0; q < nc; ++q) {
      const int nd = nodes[q];
      double o[3];
      chain_node(G, q, A, c1, c2, o);
      atomicAdd(gx + nd, o[0]);
      atomicAdd(gy + nd, o[1]);
      atomicAdd(gz + nd, o[2]);
      for (int d = 0; d < 3; ++d) {
        atomicAdd(gfi + 3 * nd + d, G.f[q] * wq * A[7 + d]);
        atomicAdd(gdfi + e * 27 + 3 * q + d, G.f[q] * wq * A[10 + d]);
      }
    }
  }
  const double s = wave_sum(nu_part);
  if (threadIdx.x == 0) atomicAdd(gnu, s);
}

// Near lists: the internal points each element integrates with the 40x40
// rule (the complement of the far test of k_svjp_far), one wave per element,
// in ascending point order.
__global__ __launch_bounds__(64) void k_near_lists(SvjpArgs a, const int* __restrict__ ptr, int* __restrict__ cnt,
                                                   int* __restrict__ list) {
  const int e = blockIdx.x;
  if (e >= a.ne) return;
  const int lane = threadIdx.x;
  const double cx = a.cen[4 * e], cy = a.cen[4 * e + 1], cz = a.cen[4 * e + 2], thr = a.cen[4 * e + 3];
  int pos = ptr ? ptr[e] : 0;
  int count = 0;
  for (int base = 0; base < a.l; base += 64) {
    const int p = base + lane;
    bool nearp = false;
    if (p < a.l) {
      const double dx = a.pts[3 * (size_t)p] - cx, dy = a.pts[3 * (size_t)p + 1] - cy,
                   dz = a.pts[3 * (size_t)p + 2] - cz;
      nearp = !(dx * dx + dy * dy + dz * dz > thr);
    }
    const unsigned long long mask = __ballot(nearp);
    if (ptr && nearp) {
      const int before = __popcll(mask & ((1ull << lane) - 1ull));
      list[pos + before] = p;
    }
    pos += __popcll(mask);
    count += __popcll(mask);
  }
  if (!ptr && lane == 0) cnt[e] = count;
}

// Near terms: one wave per (element with near points, 64 of its 1600 rule
// points); each lane sweeps the element's near points, then the wave reduces
// its contributions to the element's nodes.
__global__ __launch_bounds__(64) void k_svjp_near(SvjpArgs a, const int* __restrict__ act_elem,
                                                  const int* __restrict__ near_ptr, const int* __restrict__ near_list,
                                                  double* __restrict__ gx, double* __restrict__ gy,
                                                  double* __restrict__ gz, double* __restrict__ gfi,
                                                  double* __restrict__ gdfi, double* __restrict__ gnu) {
  __shared__ double sred[27][65];
  constexpr int kSlices = kNearPts / 64;
  const int lane = threadIdx.x;
  const int ae = blockIdx.x / kSlices;
  const int g = (blockIdx.x % kSlices) * 64 + lane;
  const int e = act_elem[ae];
  const int nc = a.nconec[e];
  const int* nodes = a.kconec0 + e * 9;
  double g1, g2, wq;
  near_rule(nc, g, g1, g2, wq);
  double acc[kAcc];
#pragma unroll
  for (int q = 0; q < kAcc; ++q) acc[q] = 0.0;
  {
    double yg[3], m[3], uw[3], tw[3];
    rule_point_state(a, e, nc, g1, g2, wq, yg, m, uw, tw);
    const int q0 = near_ptr[e], q1 = near_ptr[e + 1];
    for (int q = q0; q < q1; ++q) somi_point(a, near_list[q], yg, m, uw, tw, acc);
  }
  GpGeom G;
  gp_geometry(nc, g1, g2, nodes, a.x, a.y, a.z, G);
  double c1[3], c2[3];
  cross(G.a2, acc + 3, c1);
  cross(acc + 3, G.a1, c2);
  double part[27];
  // node coordinates
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    double o[3];
    chain_node(G, k, acc, c1, c2, o);
#pragma unroll
    for (int d = 0; d < 3; ++d) part[3 * k + d] = o[d];
  }
  scatter_nodes27(part, sred, nodes, nc, gx, gy, gz);
  // FI (node displacement) and DFI (element-node traction)
#pragma unroll
  for (int k = 0; k < 9; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) part[3 * k + d] = G.f[k] * wq * acc[7 + d];
#pragma unroll
  for (int q = 0; q < 27; ++q) sred[q][lane] = part[q];
  __syncthreads();
  if (lane < 3 * nc) {
    double s = 0.0;
    for (int j = 0; j < 64; ++j) s += sred[lane][j];
    atomicAdd(gfi + 3 * nodes[lane / 3] + lane % 3, s);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 9; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) part[3 * k + d] = G.f[k] * wq * acc[10 + d];
#pragma unroll
  for (int q = 0; q < 27; ++q) sred[q][lane] = part[q];
  __syncthreads();
  if (lane < 3 * nc) {
    double s = 0.0;
    for (int j = 0; j < 64; ++j) s += sred[lane][j];
    atomicAdd(gdfi + e * 27 + lane, s);
  }
  const double s = wave_sum(acc[6]);
  if (lane == 0) atomicAdd(gnu, s);
}

}  // namespace

// ---- host drivers -----------------------------------------------------------------
int ghmat8_vjp_run(qd_context* h, const double* lambda, const double* fi, const double* dfi, double* gx, double* gy,
                   double* gz, double* gnu) {
  const int n = h->n, ne = h->ne;
  const size_t N = 3 * (size_t)n;
  // ue: the prescribed value where kode = 0 (assembly_mod.f90:126-140), else
  // the reorganised nodal displacement
  std::vector<double> ue(27 * (size_t)ne, 0.0);
  for (int e = 0; e < ne; ++e)
    for (int ld = 0; ld < 3; ++ld)
      for (int k = 0; k < h->hnconec[e]; ++k) {
        const int nd = h->hkconec[e * 9 + k] - 1;
        ue[e * 27 + 3 * k + ld] = (h->hkode[e * 3 + ld] == 0) ? h->hbc[e * 27 + 3 * k + ld] : fi[3 * nd + ld];
      }
  std::vector<double> te(dfi, dfi + 27 * (size_t)ne);
  for (int e = 0; e < ne; ++e)
    for (int q = 3 * h->hnconec[e]; q < 27; ++q) te[e * 27 + q] = 0.0;
  DevBuf<double> dlam, duc, due, dte, drec, dshp, dg;
  QD_ALLOC(h, dlam, N);
  QD_ALLOC(h, duc, N);
  QD_ALLOC(h, due, ue.size());
  QD_ALLOC(h, dte, te.size());
  QD_ALLOC(h, drec, (size_t)ne * kGP * kRec);
  QD_ALLOC(h, dshp, 2 * (size_t)kGP * 27);
  QD_ALLOC(h, dg, N + 1);
  QD_HIP(h, hipMemcpy(dlam.p, lambda, N * sizeof(double), hipMemcpyHostToDevice));
  QD_HIP(h, hipMemcpy(duc.p, fi, N * sizeof(double), hipMemcpyHostToDevice));
  QD_HIP(h, hipMemcpy(due.p, ue.data(), ue.size() * sizeof(double), hipMemcpyHostToDevice));
  QD_HIP(h, hipMemcpy(dte.p, te.data(), te.size() * sizeof(double), hipMemcpyHostToDevice));
  QD_HIP(h, hipMemsetAsync(dg.p, 0, (N + 1) * sizeof(double), h->stream));
  AvjpArgs a{};
  a.n = n;
  a.ne = ne;
  a.mat = h->mat;
  a.x = h->x.p;
  a.y = h->y.p;
  a.z = h->z.p;
  a.nconec = h->nconec.p;
  a.kconec0 = h->kconec0.p;
  a.lam = dlam.p;
  a.uc = duc.p;
  a.ue = due.p;
  a.te = dte.p;
  double* g_x = dg.p;
  double* g_y = dg.p + n;
  double* g_z = dg.p + 2 * (size_t)n;
  double* g_nu = dg.p + N;
  hipLaunchKernelGGL(k_shape_table, dim3(1), dim3(2 * kGP), 0, h->stream, dshp.p);
  const long long slots = (long long)ne * kGP;
  hipLaunchKernelGGL(k_avjp_table, dim3((unsigned)((slots + 127) / 128)), dim3(128), 0, h->stream, a, drec.p);
  hipLaunchKernelGGL(k_avjp_regular, dim3((unsigned)((n + 63) / 64), (unsigned)((ne + kAvjpChunk - 1) / kAvjpChunk)),
                     dim3(64), 0, h->stream, a, (const double*)drec.p, (const double*)dshp.p, g_x, g_y, g_z, g_nu);
  const int npairs = (int)h->h_pairs.size();
  hipLaunchKernelGGL(k_avjp_singular, dim3((unsigned)npairs), dim3(64), 0, h->stream, a, npairs, h->pairs.p,
                     h->subtri.p, g_x, g_y, g_z, g_nu);
  QD_HIP(h, hipGetLastError());
  QD_HIP(h, hipStreamSynchronize(h->stream));
  QD_HIP(h, hipMemcpy(gx, g_x, n * sizeof(double), hipMemcpyDeviceToHost));
  QD_HIP(h, hipMemcpy(gy, g_y, n * sizeof(double), hipMemcpyDeviceToHost));
  QD_HIP(h, hipMemcpy(gz, g_z, n * sizeof(double), hipMemcpyDeviceToHost));
  QD_HIP(h, hipMemcpy(gnu, g_nu, sizeof(double), hipMemcpyDeviceToHost));
  return QD_OK;
}

int somigliana_vjp_run(qd_context* h, int l, const double* pts, const double* fi, const double* dfi, int variant,
                       const double* gu, const double* gs, double* gfi, double* gdfi, double* gx, double* gy,
                       double* gz, double* gnu) {
  const int n = h->n, ne = h->ne;
  const size_t L = (size_t)l, N = 3 * (size_t)n, E27 = 27 * (size_t)ne;
  {
    double xg[kNearGL], wg[kNearGL];
    gauss_legendre(kNearGL, xg, wg);
    QD_HIP(h, hipStreamSynchronize(h->stream));
    QD_HIP(h, hipMemcpyToSymbol(HIP_SYMBOL(d_gl40x), xg, sizeof(xg)));
    QD_HIP(h, hipMemcpyToSymbol(HIP_SYMBOL(d_gl40w), wg, sizeof(wg)));
  }
  int rc = somigliana_centres(h);
  if (rc) return rc;
  std::vector<int> fs(ne + 1), item_elem;
  int tot = 0;
  for (int e = 0; e < ne; ++e) {
    fs[e] = tot;
    const int ng = h->hnconec[e] == 9 ? 36 : 16;
    for (int k = 0; k < ng; ++k) item_elem.push_back(e);
    tot += ng;
  }
  fs[ne] = tot;
  DevBuf<double> dpts, dgu, dgs, dfi_d, ddfi, dacc, dout;
  DevBuf<int> dfs, ditem, dcnt, dptr, dlist, dact;
  QD_ALLOC(h, dpts, 3 * L);
  QD_ALLOC(h, dgu, 3 * L);
  QD_ALLOC(h, dgs, 6 * L);
  QD_ALLOC(h, dfi_d, N);
  QD_ALLOC(h, ddfi, E27);
  QD_ALLOC(h, dacc, (size_t)tot * kAcc);
  QD_ALLOC(h, dout, 3 * (size_t)n + N + E27 + 1);
  QD_ALLOC(h, dfs, fs.size());
  QD_ALLOC(h, ditem, item_elem.size());
  QD_ALLOC(h, dcnt, (size_t)ne);
  QD_ALLOC(h, dptr, (size_t)ne + 1);
  QD_HIP(h, hipMemcpy(dpts.p, pts, 3 * L * sizeof(double), hipMemcpyHostToDevice));
  if (gu) {
    QD_HIP(h, hipMemcpy(dgu.p, gu, 3 * L * sizeof(double), hipMemcpyHostToDevice));
  } else {
    QD_HIP(h, hipMemsetAsync(dgu.p, 0, 3 * L * sizeof(double), h->stream));
  }
  if (gs) {
    QD_HIP(h, hipMemcpy(dgs.p, gs, 6 * L * sizeof(double), hipMemcpyHostToDevice));
  } else {
    QD_HIP(h, hipMemsetAsync(dgs.p, 0, 6 * L * sizeof(double), h->stream));
  }
  QD_HIP(h, hipMemcpy(dfi_d.p, fi, N * sizeof(double), hipMemcpyHostToDevice));
  QD_HIP(h, hipMemcpy(ddfi.p, dfi, E27 * sizeof(double), hipMemcpyHostToDevice));
  QD_HIP(h, hipMemcpy(dfs.p, fs.data(), fs.size() * sizeof(int), hipMemcpyHostToDevice));
  QD_HIP(h, hipMemcpy(ditem.p, item_elem.data(), item_elem.size() * sizeof(int), hipMemcpyHostToDevice));
  QD_HIP(h, hipMemsetAsync(dacc.p, 0, (size_t)tot * kAcc * sizeof(double), h->stream));
  QD_HIP(h, hipMemsetAsync(dout.p, 0, (3 * (size_t)n + N + E27 + 1) * sizeof(double), h->stream));
  double* g_x = dout.p;
  double* g_y = g_x + n;
  double* g_z = g_y + n;
  double* g_fi = g_z + n;
  double* g_dfi = g_fi + N;
  double* g_nu = g_dfi + E27;

  SvjpArgs a{};
  a.l = l;
  a.ne = ne;
  a.mu = h->mat.mu;
  a.nu = h->mat.nu;
  a.sg = variant ? -1.0 : 1.0;
  a.pts = dpts.p;
  a.gu = dgu.p;
  a.gs = dgs.p;
  a.cen = h->el_centre.p;
  a.nconec = h->nconec.p;
  a.kconec0 = h->kconec0.p;
  a.x = h->x.p;
  a.y = h->y.p;
  a.z = h->z.p;
  a.fi = dfi_d.p;
  a.dfi = ddfi.p;

  const int pchunk = 256;
  hipLaunchKernelGGL(k_svjp_far, dim3((unsigned)((tot + 63) / 64), (unsigned)((l + pchunk - 1) / pchunk)), dim3(64),
                     0, h->stream, a, tot, (const int*)ditem.p, (const int*)dfs.p, pchunk, dacc.p);
  hipLaunchKernelGGL(k_svjp_far_chain, dim3((unsigned)((tot + 63) / 64)), dim3(64), 0, h->stream, a, tot,
                     (const int*)ditem.p, (const int*)dfs.p, (const double*)dacc.p, g_x, g_y, g_z, g_fi, g_dfi, g_nu);
  // near lists: counts, offsets, then the lists
  hipLaunchKernelGGL(k_near_lists, dim3((unsigned)ne), dim3(64), 0, h->stream, a, (const int*)nullptr, dcnt.p,
                     (int*)nullptr);
  QD_HIP(h, hipGetLastError());
  std::vector<int> cnt(ne), ptr(ne + 1), act;
  QD_HIP(h, hipStreamSynchronize(h->stream));
  QD_HIP(h, hipMemcpy(cnt.data(), dcnt.p, ne * sizeof(int), hipMemcpyDeviceToHost));
  ptr[0] = 0;
  for (int e = 0; e < ne; ++e) {
    ptr[e + 1] = ptr[e] + cnt[e];
    if (cnt[e] > 0) act.push_back(e);
  }
  const int nnear = ptr[ne];
  if (nnear > 0) {
    QD_ALLOC(h, dlist, (size_t)nnear);
    QD_ALLOC(h, dact, act.size());
    QD_HIP(h, hipMemcpy(dptr.p, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice));
    QD_HIP(h, hipMemcpy(dact.p, act.data(), act.size() * sizeof(int), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_near_lists, dim3((unsigned)ne), dim3(64), 0, h->stream, a, (const int*)dptr.p, dcnt.p,
                       dlist.p);
    hipLaunchKernelGGL(k_svjp_near, dim3((unsigned)(act.size() * (kNearPts / 64))), dim3(64), 0, h->stream, a,
                       (const int*)dact.p, (const int*)dptr.p, (const int*)dlist.p, g_x, g_y, g_z, g_fi, g_dfi,
                       g_nu);
  }
  QD_HIP(h, hipGetLastError());
  QD_HIP(h, hipStreamSynchronize(h->stream));
  QD_HIP(h, hipMemcpy(gx, g_x, n * sizeof(double), hipMemcpyDeviceToHost));
  QD_HIP(h, hipMemcpy(gy, g_y, n * sizeof(double), hipMemcpyDeviceToHost));
  QD_HIP(h, hipMemcpy(gz, g_z, n * sizeof(double), hipMemcpyDeviceToHost));
  QD_HIP(h, hipMemcpy(gfi, g_fi, N * sizeof(double), hipMemcpyDeviceToHost));
  QD_HIP(h, hipMemcpy(gdfi, g_dfi, E27 * sizeof(double), hipMemcpyDeviceToHost));
  QD_HIP(h, hipMemcpy(gnu, g_nu, sizeof(double), hipMemcpyDeviceToHost));
  return QD_OK;
}

}  // namespace qd
