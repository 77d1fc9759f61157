/*
 * quadesth_oracle.c -- CPU FP64 restatement of QUADESTH's hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library (tests, smoke
 * check and the benchmark's CPU-baseline leg).  Nothing in the product path
 * links or calls this file.
 *
 * Every function follows the reference routine named in its comment
 * (file:line into the QUADESTH sources) statement by statement, in the same
 * floating-point evaluation order where the reference fixes one, so that the
 * oracle and the compiled reference agree to rounding.
 *
 * Index conventions: see quadesth_oracle.h.
 */
#include "quadesth_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define QO_PI 3.14159265358979323846

/* Quadrature literals, assembly_mod.f90:6-19 (the literals are the definition;
 * the tri rule is truncated to 10 digits there and here). */
static const double gi_quad[6] = {0.238619186083197, -0.238619186083197, 0.661209386466265,
                                  -0.661209386466265, 0.932469514203152, -0.932469514203152};
static const double ome_quad[6] = {0.467913934572691, 0.467913934572691, 0.360761573048139,
                                   0.360761573048139, 0.171324492379170, 0.171324492379170};
static const double rr_tri[4] = {0.0694318422, 0.3300094782, 0.6699905218, 0.9305681558};
static const double ss_tri[4] = {0.0571041961, 0.2768430136, 0.5835904324, 0.8602401357};
static const double omei_tri[4] = {0.1739274226, 0.3260725774, 0.3260725774, 0.1739274226};
static const double omej_tri[4] = {0.1355069134, 0.2034645680, 0.1298475476, 0.0311809709};

#define H(jj, ind) h[(jj) + 3 * (ind)]
#define G(jj, ind) g[(jj) + 3 * (ind)]

/* data_mod.f90:97-108 (mu given, as the deck provides it) */
void qo_material(double mu, double nu, qo_material_t* m) {
  m->mu = mu;
  m->nu = nu;
  m->young = 2.0 * mu * (1.0 + nu);
  m->c3 = 1.0 / (16.0 * QO_PI * (1.0 - nu) * mu);
  m->c4 = 3.0 - 4.0 * nu;
  m->c5 = -1.0 / (8.0 * QO_PI * (1.0 - nu));
  m->c6 = 1.0 - 2.0 * nu;
}

/* assembly_mod.f90:247-303 == stress_mod.f90:32-112 */
void qo_shape(int ncone, double g1, double g2, double ff[9], double fd1[9], double fd2[9]) {
  for (int i = 0; i < 9; ++i) ff[i] = fd1[i] = fd2[i] = 0.0;
  if (ncone == 9) {
    ff[0] = 0.25 * g1 * g2 * (g1 - 1.0) * (g2 - 1.0);
    ff[1] = 0.5 * (1.0 - g1 * g1) * g2 * (g2 - 1.0);
    ff[2] = 0.25 * g1 * g2 * (g1 + 1.0) * (g2 - 1.0);
    ff[3] = 0.5 * g1 * (g1 + 1.0) * (1.0 - g2 * g2);
    ff[4] = 0.25 * g1 * g2 * (g1 + 1.0) * (g2 + 1.0);
    ff[5] = 0.5 * (1.0 - g1 * g1) * g2 * (g2 + 1.0);
    ff[6] = 0.25 * g1 * g2 * (g1 - 1.0) * (g2 + 1.0);
    ff[7] = 0.5 * g1 * (g1 - 1.0) * (1.0 - g2 * g2);
    ff[8] = (1.0 - g1 * g1) * (1.0 - g2 * g2);

    fd1[0] = 0.25 * g2 * (g2 - 1.0) * (2.0 * g1 - 1.0);
    fd1[1] = -g1 * g2 * (g2 - 1.0);
    fd1[2] = 0.25 * g2 * (g2 - 1.0) * (2.0 * g1 + 1.0);
    fd1[3] = 0.5 * (1.0 - g2 * g2) * (2.0 * g1 + 1.0);
    fd1[4] = 0.25 * g2 * (g2 + 1.0) * (2.0 * g1 + 1.0);
    fd1[5] = -g1 * g2 * (g2 + 1.0);
    fd1[6] = 0.25 * g2 * (g2 + 1.0) * (2.0 * g1 - 1.0);
    fd1[7] = 0.5 * (1.0 - g2 * g2) * (2.0 * g1 - 1.0);
    fd1[8] = -2.0 * g1 * (1.0 - g2 * g2);

    fd2[0] = 0.25 * g1 * (g1 - 1.0) * (2.0 * g2 - 1.0);
    fd2[1] = 0.5 * (1.0 - g1 * g1) * (2.0 * g2 - 1.0);
    fd2[2] = 0.25 * g1 * (g1 + 1.0) * (2.0 * g2 - 1.0);
    fd2[3] = -g1 * g2 * (g1 + 1.0);
    fd2[4] = 0.25 * g1 * (g1 + 1.0) * (2.0 * g2 + 1.0);
    fd2[5] = 0.5 * (1.0 - g1 * g1) * (2.0 * g2 + 1.0);
    fd2[6] = 0.25 * g1 * (g1 - 1.0) * (2.0 * g2 + 1.0);
    fd2[7] = -g1 * g2 * (g1 - 1.0);
    fd2[8] = -2.0 * g2 * (1.0 - g1 * g1);
  } else {
    double g3 = 1.0 - g1 - g2;
    ff[0] = g1 * (2.0 * g1 - 1.0);
    ff[1] = g2 * (2.0 * g2 - 1.0);
    ff[2] = g3 * (2.0 * g3 - 1.0);
    ff[3] = 4.0 * g1 * g2;
    ff[4] = 4.0 * g2 * g3;
    ff[5] = 4.0 * g3 * g1;

    fd1[0] = 4.0 * g1 - 1.0;
    fd1[1] = 0.0;
    fd1[2] = 1.0 - 4.0 * g3;
    fd1[3] = 4.0 * g2;
    fd1[4] = -4.0 * g2;
    fd1[5] = 4.0 * (g3 - g1);

    fd2[0] = 0.0;
    fd2[1] = 4.0 * g2 - 1.0;
    fd2[2] = 1.0 - 4.0 * g3;
    fd2[3] = 4.0 * g1;
    fd2[4] = 4.0 * (g3 - g2);
    fd2[5] = -4.0 * g1;
  }
}

/* Unnormalised normal (s1, s2, s3) from intrinsic derivatives, in the
 * reference's dc/dct double-loop form (assembly_mod.f90:311-329). */
static void normal_from_derivs(int ncone, const double fd1[9], const double fd2[9],
                               const double xj1[9], const double xj2[9], const double xj3[9],
                               double* s1, double* s2, double* s3) {
  double aux[9], bux[9], cux[9];
  for (int jk = 0; jk < ncone; ++jk) {
    aux[jk] = bux[jk] = cux[jk] = 0.0;
    for (int ik = 0; ik < ncone; ++ik) {
      double d = fd1[ik] * fd2[jk] - fd2[ik] * fd1[jk];
      aux[jk] += xj2[ik] * d;
      bux[jk] += xj3[ik] * d;
      cux[jk] += xj1[ik] * d;
    }
  }
  double a = 0.0, b = 0.0, c = 0.0;
  for (int jk = 0; jk < ncone; ++jk) {
    a += aux[jk] * xj3[jk];
    b += bux[jk] * xj1[jk];
    c += cux[jk] * xj2[jk];
  }
  *s1 = a;
  *s2 = b;
  *s3 = c;
}

/* assembly_mod.f90:225-337 */
void qo_etaxja(const double xj1[9], const double xj2[9], const double xj3[9],
               double eta1[36], double eta2[36], double eta3[36], double xja[36],
               double xg1[36], double xg2[36], double xg3[36], double ff[9 * 36], int ncone) {
  int npg = (ncone < 9) ? 4 : 6;
  for (int k = 0; k < 36; ++k) {
    eta1[k] = eta2[k] = eta3[k] = xja[k] = xg1[k] = xg2[k] = xg3[k] = 0.0;
    for (int m = 0; m < 9; ++m) ff[k * 9 + m] = 0.0;
  }
  for (int i = 0; i < npg; ++i) {
    for (int j = 0; j < npg; ++j) {
      int k = npg * i + j;
      double g1, g2, f[9], fd1[9], fd2[9];
      if (ncone == 9) {
        g1 = gi_quad[i];
        g2 = gi_quad[j];
      } else {
        g1 = ss_tri[j];
        g2 = rr_tri[i] * (1.0 - ss_tri[j]);
      }
      qo_shape(ncone, g1, g2, f, fd1, fd2);
      for (int m = 0; m < ncone; ++m) ff[k * 9 + m] = f[m];
      for (int ik = 0; ik < ncone; ++ik) {
        xg1[k] += f[ik] * xj1[ik];
        xg2[k] += f[ik] * xj2[ik];
        xg3[k] += f[ik] * xj3[ik];
      }
      double s1, s2, s3;
      normal_from_derivs(ncone, fd1, fd2, xj1, xj2, xj3, &s1, &s2, &s3);
      xja[k] = sqrt(s1 * s1 + s2 * s2 + s3 * s3);
      eta1[k] = s1 / xja[k];
      eta2[k] = s2 / xja[k];
      eta3[k] = s3 / xja[k];
    }
  }
}

static void gather_coords(int iej, const int* nconec, const int* kconec, const double* x,
                          const double* y, const double* z, double xj1[9], double xj2[9],
                          double xj3[9]) {
  for (int m = 0; m < 9; ++m) xj1[m] = xj2[m] = xj3[m] = 0.0;
  for (int m = 0; m < nconec[iej]; ++m) {
    int k = kconec[iej * 9 + m] - 1;
    xj1[m] = x[k];
    xj2[m] = y[k];
    xj3[m] = z[k];
  }
}

/* assembly_mod.f90:28-51 */
void qo_precompute(int ne, const int* nconec, const int* kconec, const double* x,
                   const double* y, const double* z, double* eta1, double* eta2, double* eta3,
                   double* xja, double* xg1, double* xg2, double* xg3, double* ff) {
#pragma omp parallel for schedule(static)
  for (int e = 0; e < ne; ++e) {
    double xj1[9], xj2[9], xj3[9];
    gather_coords(e, nconec, kconec, x, y, z, xj1, xj2, xj3);
    qo_etaxja(xj1, xj2, xj3, eta1 + 36 * e, eta2 + 36 * e, eta3 + 36 * e, xja + 36 * e,
              xg1 + 36 * e, xg2 + 36 * e, xg3 + 36 * e, ff + 36 * 9 * e, nconec[e]);
  }
}

/* assembly_mod.f90:618-650; u(jj,ii) at u[jj + 3*ii] */
void qo_funkelvin(const qo_material_t* m, double xp, double yp, double zp, const double eta[3],
                  double xg, double yg, double zg, double u[9], double t[9]) {
  double r1 = xg - xp, r2 = yg - yp, r3 = zg - zp;
  double r = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
  double rdn = (r1 * eta[0] + r2 * eta[1] + r3 * eta[2]) / r;
  double rd[3] = {r1 / r, r2 / r, r3 / r};
  double c3r = m->c3 / r;
  double c5r2 = m->c5 / (r * r);
  for (int ii = 0; ii < 3; ++ii) {
    for (int jj = 0; jj < 3; ++jj) {
      double dl = (ii == jj) ? 1.0 : 0.0;
      u[jj + 3 * ii] = (m->c4 * dl + rd[ii] * rd[jj]) * c3r;
      t[jj + 3 * ii] = c5r2 * (rdn * (m->c6 * dl + 3.0 * rd[ii] * rd[jj]) +
                               m->c6 * (eta[jj] * rd[ii] - eta[ii] * rd[jj]));
    }
  }
}

/* assembly_mod.f90:339-384 */
void qo_extin8(const qo_material_t* m, const double* eta1, const double* eta2, const double* eta3,
               const double* xja, const double* xg1, const double* xg2, const double* xg3,
               const double* ff, const double xi[3], double h[81], double g[81], int ncone) {
  for (int i = 0; i < 81; ++i) h[i] = g[i] = 0.0;
  int npg = (ncone < 9) ? 4 : 6;
  for (int i = 0; i < npg; ++i) {
    for (int j = 0; j < npg; ++j) {
      int k = npg * i + j;
      double eta[3] = {eta1[k], eta2[k], eta3[k]};
      double p12[9];
      for (int ik = 0; ik < ncone; ++ik) {
        if (ncone == 9)
          p12[ik] = ff[k * 9 + ik] * xja[k] * ome_quad[i] * ome_quad[j];
        else
          p12[ik] = ff[k * 9 + ik] * xja[k] * omei_tri[i] * omej_tri[j];
      }
      double u[9], t[9];
      qo_funkelvin(m, xi[0], xi[1], xi[2], eta, xg1[k], xg2[k], xg3[k], u, t);
      for (int kk = 0; kk < ncone; ++kk)
        for (int ii = 0; ii < 3; ++ii)
          for (int jj = 0; jj < 3; ++jj) {
            int ind = 3 * kk + ii;
            G(jj, ind) += u[jj + 3 * ii] * p12[kk];
            H(jj, ind) += t[jj + 3 * ii] * p12[kk];
          }
    }
  }
}

/* assembly_mod.f90:492-616 */
static void qo_triloc8(const qo_material_t* m, const double xi[3], const double xj1[9],
                       const double xj2[9], const double xj3[9], double h[81], double g[81],
                       double area, const double v1[3], const double v2[3], int ncone) {
  for (int i = 0; i < 6; ++i) {
    for (int j = 0; j < 6; ++j) {
      double ss1 = 0.5 * (gi_quad[i] + 1.0);
      double ss2 = 0.5 * (gi_quad[j] + 1.0);
      double g1 = (1.0 - ss1) * v1[0] + ss1 * (1.0 - ss2) * v1[1] + ss1 * ss2 * v1[2];
      double g2 = (1.0 - ss1) * v2[0] + ss1 * (1.0 - ss2) * v2[1] + ss1 * ss2 * v2[2];
      double f[9], fd1[9], fd2[9];
      qo_shape(ncone, g1, g2, f, fd1, fd2);
      double s1, s2, s3;
      normal_from_derivs(ncone, fd1, fd2, xj1, xj2, xj3, &s1, &s2, &s3);
      double xja = sqrt(s1 * s1 + s2 * s2 + s3 * s3);
      double eta[3] = {s1 / xja, s2 / xja, s3 / xja};
      double xp[3] = {0.0, 0.0, 0.0};
      for (int ik = 0; ik < ncone; ++ik) {
        xp[0] += f[ik] * xj1[ik];
        xp[1] += f[ik] * xj2[ik];
        xp[2] += f[ik] * xj3[ik];
      }
      double p12[9];
      for (int ik = 0; ik < ncone; ++ik)
        p12[ik] = f[ik] * xja * (gi_quad[i] + 1.0) * area * 0.25 * ome_quad[i] * ome_quad[j];
      double u[9], t[9];
      qo_funkelvin(m, xi[0], xi[1], xi[2], eta, xp[0], xp[1], xp[2], u, t);
      for (int kk = 0; kk < ncone; ++kk)
        for (int ii = 0; ii < 3; ++ii)
          for (int jj = 0; jj < 3; ++jj) {
            int ind = 3 * kk + ii;
            G(jj, ind) += u[jj + 3 * ii] * p12[kk];
            H(jj, ind) += t[jj + 3 * ii] * p12[kk];
          }
    }
  }
}

/* assembly_mod.f90:386-490; nodo is 1-based as in the reference */
void qo_locin8(const qo_material_t* m, const double xj1[9], const double xj2[9],
               const double xj3[9], double h[81], double g[81], int nodo, int ncone) {
  for (int i = 0; i < 81; ++i) h[i] = g[i] = 0.0;
  double xi[3] = {xj1[nodo - 1], xj2[nodo - 1], xj3[nodo - 1]};
  double v1[3], v2[3], area;
  if (ncone == 9) {
    static const double gi1[9] = {-1.0, 0.0, 1.0, 1.0, 1.0, 0.0, -1.0, -1.0, 0.0};
    static const double gi2[9] = {-1.0, -1.0, -1.0, 0.0, 1.0, 1.0, 1.0, 0.0, 0.0};
    double gii1[16], gii2[16]; /* 1-based, entries 1..15 */
    for (int i = 1; i <= 8; ++i) {
      gii1[i] = gi1[i - 1];
      gii2[i] = gi2[i - 1];
    }
    for (int i = 9; i <= 15; ++i) {
      gii1[i] = gi1[i - 9];
      gii2[i] = gi2[i - 9];
    }
    v1[0] = gi1[nodo - 1];
    v2[0] = gi2[nodo - 1];
    if (nodo == 1 || nodo == 3 || nodo == 5 || nodo == 7) {
      area = 2.0;
      for (int k = 1; k <= 2; ++k) {
        int mm = 2 * k;
        v1[1] = gii1[nodo + mm];
        v2[1] = gii2[nodo + mm];
        v1[2] = gii1[nodo + mm + 2];
        v2[2] = gii2[nodo + mm + 2];
        qo_triloc8(m, xi, xj1, xj2, xj3, h, g, area, v1, v2, 9);
      }
    } else if (nodo == 2 || nodo == 4 || nodo == 6 || nodo == 8) {
      for (int k = 1; k <= 3; ++k) {
        int mm = 2 * k - 1;
        v1[1] = gii1[nodo + mm];
        v2[1] = gii2[nodo + mm];
        v1[2] = gii1[nodo + mm + 2];
        v2[2] = gii2[nodo + mm + 2];
        area = 0.5 * (v1[1] * v2[2] + v1[0] * v2[1] + v1[2] * v2[0] - v1[1] * v2[0] -
                      v1[2] * v2[1] - v1[0] * v2[2]);
        qo_triloc8(m, xi, xj1, xj2, xj3, h, g, area, v1, v2, 9);
      }
    } else {
      area = 1.0;
      for (int k = 1; k <= 4; ++k) {
        int mm = 2 * k - 1;
        v1[1] = gii1[mm];
        v2[1] = gii2[mm];
        v1[2] = gii1[mm + 2];
        v2[2] = gii2[mm + 2];
        qo_triloc8(m, xi, xj1, xj2, xj3, h, g, area, v1, v2, 9);
      }
    }
  } else {
    static const double gi1[6] = {1.0, 0.0, 0.0, 0.5, 0.0, 0.5};
    static const double gi2[6] = {0.0, 1.0, 0.0, 0.5, 0.5, 0.0};
    v1[0] = gi1[nodo - 1];
    v2[0] = gi2[nodo - 1];
    /* sub-triangles (second, third vertex as 1-based tri node numbers) */
    int sub[2][2];
    int nsub;
    switch (nodo) {
      case 1: area = 0.5; nsub = 1; sub[0][0] = 2; sub[0][1] = 3; break;
      case 2: area = 0.5; nsub = 1; sub[0][0] = 3; sub[0][1] = 1; break;
      case 3: area = 0.5; nsub = 1; sub[0][0] = 1; sub[0][1] = 2; break;
      case 4: area = 0.25; nsub = 2; sub[0][0] = 2; sub[0][1] = 3; sub[1][0] = 3; sub[1][1] = 1; break;
      case 5: area = 0.25; nsub = 2; sub[0][0] = 3; sub[0][1] = 1; sub[1][0] = 1; sub[1][1] = 2; break;
      default: area = 0.25; nsub = 2; sub[0][0] = 1; sub[0][1] = 2; sub[1][0] = 2; sub[1][1] = 3; break;
    }
    for (int s = 0; s < nsub; ++s) {
      v1[1] = gi1[sub[s][0] - 1];
      v2[1] = gi2[sub[s][0] - 1];
      v1[2] = gi1[sub[s][1] - 1];
      v2[2] = gi2[sub[s][1] - 1];
      qo_triloc8(m, xi, xj1, xj2, xj3, h, g, area, v1, v2, 6);
    }
  }
}

/* Entries per node of ke.  8 is the reference's ke(8, n) (data_mod.f90:79,
 * io_mod.f90:131); a larger width lets the kode0 search and the
 * reorganisation see every element of a node shared by more than 8, where
 * the reference reads past ke(8, .) (assembly_mod.f90:150-151,
 * io_mod.f90:170-171). */
static int g_kecap = 8;
void qo_set_ke_cap(int cap) { g_kecap = cap > 8 ? cap : 8; }
int qo_ke_cap(void) { return g_kecap; }

/* io_mod.f90:124-135 (ke holds the first g_kecap elements of each node) */
void qo_adjacency(int n, int ne, const int* nconec, const int* kconec, int* nesim, int* ke) {
  for (int i = 0; i < n; ++i) nesim[i] = 0;
  for (long i = 0; i < (long)g_kecap * n; ++i) ke[i] = 0;
  for (int e = 0; e < ne; ++e)
    for (int j = 0; j < nconec[e]; ++j) {
      int node = kconec[e * 9 + j] - 1;
      nesim[node] += 1;
      if (nesim[node] <= g_kecap) ke[(long)node * g_kecap + nesim[node] - 1] = e + 1;
    }
}

/* kode0 / vp search of assembly_mod.f90:149-161 (and 194-206): first element
 * of node ik (ke order) prescribing the displacement in direction ld. The
 * reference loops to nesim(ik) (which overflows ke when nesim > 8); the
 * search here stops at the stored entries. */
static int find_vp(int ik, int ld, const int* nconec, const int* kconec, const int* kode,
                   const double* bc, const int* nesim, const int* ke, double* vp) {
  int lim = nesim[ik] < g_kecap ? nesim[ik] : g_kecap;
  for (int i = 0; i < lim; ++i) {
    int e = ke[(long)ik * g_kecap + i] - 1;
    if (kode[e * 3 + ld] == 0) {
      for (int mm = 0; mm < nconec[e]; ++mm)
        if (kconec[e * 9 + mm] - 1 == ik) {
          *vp = bc[e * 27 + 3 * mm + ld];
          return 1;
        }
    }
  }
  return 0;
}

void qo_element_pair(int in, int iej, const double* x, const double* y, const double* z,
                     const int* nconec, const int* kconec, const qo_material_t* m,
                     const double* eta1, const double* eta2, const double* eta3, const double* xja,
                     const double* xg1, const double* xg2, const double* xg3, const double* ff,
                     double h[81], double g[81]) {
  double xj1[9], xj2[9], xj3[9];
  gather_coords(iej, nconec, kconec, x, y, z, xj1, xj2, xj3);
  int nodo = 0;
  for (int i = 0; i < nconec[iej]; ++i)
    if (in == kconec[iej * 9 + i] - 1) {
      nodo = i + 1;
      break;
    }
  if (nodo > 0) {
    qo_locin8(m, xj1, xj2, xj3, h, g, nodo, nconec[iej]);
  } else {
    double xi[3] = {x[in], y[in], z[in]};
    qo_extin8(m, eta1 + 36 * iej, eta2 + 36 * iej, eta3 + 36 * iej, xja + 36 * iej,
              xg1 + 36 * iej, xg2 + 36 * iej, xg3 + 36 * iej, ff + 36 * 9 * iej, xi, h, g,
              nconec[iej]);
  }
}

void qo_ghmat8_node(int in, int n, int ne, const double* x, const double* y, const double* z,
                    const int* nconec, const int* kconec, const int* kode, const double* bc,
                    const int* nesim, const int* ke, const qo_material_t* m,
                    const double* eta1, const double* eta2, const double* eta3, const double* xja,
                    const double* xg1, const double* xg2, const double* xg3, const double* ff,
                    double* A, long long ldA, int r0, double* FIr);

/* assembly_mod.f90:53-223 */
void qo_ghmat8(int n, int ne, const double* x, const double* y, const double* z,
               const int* nconec, const int* kconec, const int* kode, const double* bc,
               const int* nesim, const int* ke, const qo_material_t* m,
               const double* eta1, const double* eta2, const double* eta3, const double* xja,
               const double* xg1, const double* xg2, const double* xg3, const double* ff,
               double* A, double* FI) {
  const long long N = 3LL * n;
  memset(A, 0, sizeof(double) * (size_t)(N * N));
  memset(FI, 0, sizeof(double) * (size_t)N);
#pragma omp parallel for schedule(dynamic)
  for (int in = 0; in < n; ++in)
    qo_ghmat8_node(in, n, ne, x, y, z, nconec, kconec, kode, bc, nesim, ke, m, eta1, eta2, eta3, xja,
                   xg1, xg2, xg3, ff, A, N, 3 * in, FI + 3 * in);
}

/* Rows 3in..3in+2 of A (written at A[col*ldA + r0 + jj]) and of FI for one
 * collocation node: the body of ghmat8's node loop (assembly_mod.f90:72-220). */
void qo_ghmat8_node(int in, int n, int ne, const double* x, const double* y, const double* z,
                    const int* nconec, const int* kconec, const int* kode, const double* bc,
                    const int* nesim, const int* ke, const qo_material_t* m,
                    const double* eta1, const double* eta2, const double* eta3, const double* xja,
                    const double* xg1, const double* xg2, const double* xg3, const double* ff,
                    double* A, long long ldA, int r0, double* FIr) {
  (void)n;
  const double rge = m->mu;
  double* FI = FIr - r0;
#define AA(r, c) A[(long long)(c) * ldA + (r)]
  {
    double h[81], g[81];
    for (int iej = 0; iej < ne; ++iej) {
      int nc = nconec[iej];
      int nodo = 0;
      for (int i = 0; i < nc; ++i)
        if (in == kconec[iej * 9 + i] - 1) {
          nodo = i + 1;
          break;
        }
      qo_element_pair(in, iej, x, y, z, nconec, kconec, m, eta1, eta2, eta3, xja, xg1, xg2, xg3,
                      ff, h, g);
      double d1[9] = {0}, d2[9] = {0}; /* d(i,j) at [i + 3j] */
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i)
          for (int k = 0; k < nc; ++k) {
            int ind = 3 * k + j;
            if (nodo == 0)
              d1[i + 3 * j] -= H(i, ind);
            else if (k != nodo - 1)
              d2[i + 3 * j] -= H(i, ind);
          }
      for (int ld = 0; ld < 3; ++ld) {
        if (kode[iej * 3 + ld] == 0) {
          for (int k = 0; k < nc; ++k) {
            int ik = kconec[iej * 9 + k] - 1;
            int col = 3 * ik + ld;
            int ind = 3 * k + ld;
            double bcv = bc[iej * 27 + 3 * k + ld];
            for (int jj = 0; jj < 3; ++jj) AA(r0 + jj, col) -= G(jj, ind) * rge;
            if (in != ik) {
              for (int jj = 0; jj < 3; ++jj) FI[r0 + jj] -= H(jj, ind) * bcv;
            } else {
              for (int jj = 0; jj < 3; ++jj) FI[r0 + jj] -= d2[jj + 3 * ld] * bcv;
            }
          }
        } else {
          for (int k = 0; k < nc; ++k) {
            int ik = kconec[iej * 9 + k] - 1;
            int ind = 3 * k + ld;
            double bcv = bc[iej * 27 + 3 * k + ld];
            double vp = 0.0;
            int kode0 = find_vp(ik, ld, nconec, kconec, kode, bc, nesim, ke, &vp);
            if (kode0 == 0) {
              for (int jj = 0; jj < 3; ++jj) FI[r0 + jj] += G(jj, ind) * bcv;
              if (in != ik) {
                for (int jj = 0; jj < 3; ++jj) AA(r0 + jj, 3 * ik + ld) += H(jj, ind);
              } else {
                for (int jj = 0; jj < 3; ++jj) AA(r0 + jj, 3 * in + ld) += d2[jj + 3 * ld];
              }
            } else {
              if (in != ik) {
                for (int jj = 0; jj < 3; ++jj) FI[r0 + jj] = FI[r0 + jj] + G(jj, ind) * bcv - H(jj, ind) * vp;
              } else {
                for (int jj = 0; jj < 3; ++jj) FI[r0 + jj] = FI[r0 + jj] + G(jj, ind) * bcv - d2[jj + 3 * ld] * vp;
              }
            }
          }
        }
      }
      if (nodo == 0) {
        for (int ld = 0; ld < 3; ++ld) {
          double vp = 0.0;
          int kode0 = find_vp(in, ld, nconec, kconec, kode, bc, nesim, ke, &vp);
          if (kode0 == 0) {
            for (int jj = 0; jj < 3; ++jj) AA(r0 + jj, 3 * in + ld) += d1[jj + 3 * ld];
          } else {
            for (int jj = 0; jj < 3; ++jj) FI[r0 + jj] -= d1[jj + 3 * ld] * vp;
          }
        }
      }
    }
  }
#undef AA
}

/* Selected collocation nodes only: Arows[(i*3+jj)*N + col] (row-major rows),
 * FIrows[i*3+jj]. */
void qo_ghmat8_rows(int n, int ne, const double* x, const double* y, const double* z,
                    const int* nconec, const int* kconec, const int* kode, const double* bc,
                    const int* nesim, const int* ke, const qo_material_t* m,
                    const double* eta1, const double* eta2, const double* eta3, const double* xja,
                    const double* xg1, const double* xg2, const double* xg3, const double* ff,
                    int nsel, const int* nodes, double* Arows, double* FIrows) {
  const long long N = 3LL * n;
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < nsel; ++i) {
    double* tmp = (double*)calloc((size_t)(3 * N), sizeof(double));
    double f[3] = {0, 0, 0};
    qo_ghmat8_node(nodes[i], n, ne, x, y, z, nconec, kconec, kode, bc, nesim, ke, m, eta1, eta2, eta3,
                   xja, xg1, xg2, xg3, ff, tmp, 3, 0, f);
    for (int jj = 0; jj < 3; ++jj) {
      for (long long c = 0; c < N; ++c) Arows[((long long)i * 3 + jj) * N + c] = tmp[c * 3 + jj];
      FIrows[i * 3 + jj] = f[jj];
    }
    free(tmp);
  }
}

/* DGESV as LAPACK defines it (dgetrf: partial pivoting with the first
 * largest |a| as pivot, multipliers by the reciprocal pivot; dgetrs: row
 * interchanges on B, unit-lower then upper substitution).  Column-major,
 * 1-based ipiv.  Returns info (0, or the first zero pivot, 1-based). */
int qo_dgesv(int n, int nrhs, double* a, int lda, int* ipiv, double* b, int ldb) {
  int info = 0;
#define AM(i, j) a[(long long)(j) * lda + (i)]
  for (int j = 0; j < n; ++j) {
    int p = j;
    double best = fabs(AM(j, j));
    for (int i = j + 1; i < n; ++i) {
      double v = fabs(AM(i, j));
      if (v > best) {
        best = v;
        p = i;
      }
    }
    ipiv[j] = p + 1;
    if (AM(p, j) != 0.0) {
      if (p != j)
        for (int c = 0; c < n; ++c) {
          double t = AM(j, c);
          AM(j, c) = AM(p, c);
          AM(p, c) = t;
        }
      double rp = 1.0 / AM(j, j);
      for (int i = j + 1; i < n; ++i) AM(i, j) *= rp;
    } else if (info == 0) {
      info = j + 1;
    }
#pragma omp parallel for schedule(static)
    for (int c = j + 1; c < n; ++c) {
      double ujc = AM(j, c);
      if (ujc != 0.0)
        for (int i = j + 1; i < n; ++i) AM(i, c) -= AM(i, j) * ujc;
    }
  }
  if (info != 0) return info;
  for (int r = 0; r < nrhs; ++r) {
    double* bb = b + (long long)r * ldb;
    for (int j = 0; j < n; ++j) {
      int p = ipiv[j] - 1;
      if (p != j) {
        double t = bb[j];
        bb[j] = bb[p];
        bb[p] = t;
      }
    }
    for (int j = 0; j < n; ++j) {
      double v = bb[j];
      if (v != 0.0)
        for (int i = j + 1; i < n; ++i) bb[i] -= v * AM(i, j);
    }
    for (int j = n - 1; j >= 0; --j) {
      if (bb[j] != 0.0) {
        bb[j] /= AM(j, j);
        double v = bb[j];
        for (int i = 0; i < j; ++i) bb[i] -= v * AM(i, j);
      }
    }
  }
#undef AM
  return 0;
}

/* io_mod.f90:148-183 (node loop bounded by the stored ke entries) */
void qo_reorganize(int n, int ne, const int* nconec, const int* kconec, const int* kode,
                   const double* bc, const int* nesim, const int* ke, double mu, double* FI,
                   double* DFI) {
  for (int ie = 0; ie < ne; ++ie)
    for (int ld = 0; ld < 3; ++ld) {
      if (kode[ie * 3 + ld] == 0) {
        for (int k = 0; k < nconec[ie]; ++k) {
          int ik = kconec[ie * 9 + k] - 1;
          DFI[ie * 27 + 3 * k + ld] = FI[3 * ik + ld] * mu;
        }
      } else {
        for (int k = 0; k < nconec[ie]; ++k) DFI[ie * 27 + 3 * k + ld] = bc[ie * 27 + 3 * k + ld];
      }
    }
  for (int in = 0; in < n; ++in)
    for (int ld = 0; ld < 3; ++ld) {
      int lim = nesim[in] < g_kecap ? nesim[in] : g_kecap;
      for (int j = 0; j < lim; ++j) {
        int e = ke[(long)in * g_kecap + j] - 1;
        if (kode[e * 3 + ld] == 0) {
          for (int mm = 0; mm < nconec[e]; ++mm)
            if (kconec[e * 9 + mm] - 1 == in) {
              FI[3 * in + ld] = bc[e * 27 + 3 * mm + ld];
              break;
            }
          break;
        }
      }
    }
}

/* stress_mod.f90:475-503 */
void qo_gauss_legendre(int npts, double* xg, double* wg) {
  int m_half = (npts + 1) / 2;
  for (int i = 1; i <= m_half; ++i) {
    double z = cos(QO_PI * ((double)i - 0.25) / ((double)npts + 0.5));
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

/* stress_mod.f90:299-359 */
void qo_principal_vm(const double sig[6], double prin[3], double* vm) {
  const double tol = 1.0e-20;
  double sxx = sig[0], syy = sig[1], szz = sig[2], sxy = sig[3], syz = sig[4], szx = sig[5];
  *vm = sqrt(0.5 * ((sxx - syy) * (sxx - syy) + (syy - szz) * (syy - szz) +
                    (szz - sxx) * (szz - sxx) + 6.0 * (sxy * sxy + syz * syz + szx * szx)));
  double i1 = sxx + syy + szz;
  double q = i1 / 3.0;
  double pxx = sxx - q, pyy = syy - q, pzz = szz - q;
  double j2 = 0.5 * (pxx * pxx + pyy * pyy + pzz * pzz) + sxy * sxy + syz * syz + szx * szx;
  if (j2 <= tol * (i1 * i1 + 1.0)) {
    prin[0] = prin[1] = prin[2] = q;
    return;
  }
  double j3 = pxx * (pyy * pzz - syz * syz) - sxy * (sxy * pzz - syz * szx) +
              szx * (sxy * syz - pyy * szx);
  double r = sqrt(4.0 * j2 / 3.0);
  double c3p = 4.0 * j3 / (r * r * r);
  if (c3p > 1.0) c3p = 1.0;
  if (c3p < -1.0) c3p = -1.0;
  double phi = acos(c3p) / 3.0;
  prin[0] = q + r * cos(phi);
  prin[2] = q + r * cos(phi + 2.0 * QO_PI / 3.0);
  prin[1] = 3.0 * q - prin[0] - prin[2];
  double tmp;
  if (prin[0] < prin[1]) { tmp = prin[0]; prin[0] = prin[1]; prin[1] = tmp; }
  if (prin[1] < prin[2]) { tmp = prin[1]; prin[1] = prin[2]; prin[2] = tmp; }
  if (prin[0] < prin[1]) { tmp = prin[0]; prin[0] = prin[1]; prin[1] = tmp; }
}

/* One pre-weighted Gauss point of the Somigliana tables. */
typedef struct {
  double x, y, z, nx, ny, nz, u[3], t[3];
} qo_gp_t;

/* stress_mod.f90:788-886 for one Gauss point */
static void somigliana_gp(const qo_gp_t* gp, double xi_x, double xi_y, double xi_z, double nu,
                          double c4, double c6, double pr4, double cU, double cT, double cD,
                          double cS, int variant, double u_acc[3], double sig_acc[6]) {
  double rx = gp->x - xi_x, ry = gp->y - xi_y, rz = gp->z - xi_z;
  double r2 = rx * rx + ry * ry + rz * rz;
  double r = sqrt(r2);
  double inv_r = 1.0 / r;
  double inv_r2 = inv_r * inv_r;
  double inv_r3 = inv_r2 * inv_r;
  double rd1 = rx * inv_r, rd2 = ry * inv_r, rd3 = rz * inv_r;
  double n1 = gp->nx, n2 = gp->ny, n3 = gp->nz;
  double rdn = rd1 * n1 + rd2 * n2 + rd3 * n3;
  double tu1 = gp->u[0], tu2 = gp->u[1], tu3 = gp->u[2];
  double tt1 = gp->t[0], tt2 = gp->t[1], tt3 = gp->t[2];
  double rdt = rd1 * tt1 + rd2 * tt2 + rd3 * tt3;
  double rdu = rd1 * tu1 + rd2 * tu2 + rd3 * tu3;
  double nuv = n1 * tu1 + n2 * tu2 + n3 * tu3;
  double fU = cU * inv_r, fT = cT * inv_r2;
  /* variant 1 flips the antisymmetric term to the sign the assembly kernel
   * uses (assembly_mod.f90:646-647) */
  double sg = variant ? -1.0 : 1.0;
  u_acc[0] += fU * (c4 * tt1 + rd1 * rdt) -
              fT * (rdn * (c6 * tu1 + 3.0 * rd1 * rdu) + sg * c6 * (nuv * rd1 - n1 * rdu));
  u_acc[1] += fU * (c4 * tt2 + rd2 * rdt) -
              fT * (rdn * (c6 * tu2 + 3.0 * rd2 * rdu) + sg * c6 * (nuv * rd2 - n2 * rdu));
  u_acc[2] += fU * (c4 * tt3 + rd3 * rdt) -
              fT * (rdn * (c6 * tu3 + 3.0 * rd3 * rdu) + sg * c6 * (nuv * rd3 - n3 * rdu));
  double fD = cD * inv_r2, fS = cS * inv_r3;
  double Dxx = fD * (c6 * (2.0 * tt1 * rd1 - rdt) + 3.0 * rd1 * rd1 * rdt);
  double Dyy = fD * (c6 * (2.0 * tt2 * rd2 - rdt) + 3.0 * rd2 * rd2 * rdt);
  double Dzz = fD * (c6 * (2.0 * tt3 * rd3 - rdt) + 3.0 * rd3 * rd3 * rdt);
  double Dxy = fD * (c6 * (tt1 * rd2 + tt2 * rd1) + 3.0 * rd1 * rd2 * rdt);
  double Dyz = fD * (c6 * (tt2 * rd3 + tt3 * rd2) + 3.0 * rd2 * rd3 * rdt);
  double Dzx = fD * (c6 * (tt3 * rd1 + tt1 * rd3) + 3.0 * rd3 * rd1 * rdt);
  double Sxx = fS * (3.0 * rdn * (c6 * rdu + 2.0 * nu * tu1 * rd1 - 5.0 * rd1 * rd1 * rdu) +
                     6.0 * nu * rdu * n1 * rd1 + c6 * (3.0 * nuv * rd1 * rd1 + 2.0 * n1 * tu1) -
                     pr4 * nuv);
  double Syy = fS * (3.0 * rdn * (c6 * rdu + 2.0 * nu * tu2 * rd2 - 5.0 * rd2 * rd2 * rdu) +
                     6.0 * nu * rdu * n2 * rd2 + c6 * (3.0 * nuv * rd2 * rd2 + 2.0 * n2 * tu2) -
                     pr4 * nuv);
  double Szz = fS * (3.0 * rdn * (c6 * rdu + 2.0 * nu * tu3 * rd3 - 5.0 * rd3 * rd3 * rdu) +
                     6.0 * nu * rdu * n3 * rd3 + c6 * (3.0 * nuv * rd3 * rd3 + 2.0 * n3 * tu3) -
                     pr4 * nuv);
  double Sxy = fS * (3.0 * rdn * (nu * (tu1 * rd2 + tu2 * rd1) - 5.0 * rd1 * rd2 * rdu) +
                     3.0 * nu * rdu * (n1 * rd2 + n2 * rd1) +
                     c6 * (3.0 * nuv * rd1 * rd2 + n1 * tu2 + n2 * tu1));
  double Syz = fS * (3.0 * rdn * (nu * (tu2 * rd3 + tu3 * rd2) - 5.0 * rd2 * rd3 * rdu) +
                     3.0 * nu * rdu * (n2 * rd3 + n3 * rd2) +
                     c6 * (3.0 * nuv * rd2 * rd3 + n2 * tu3 + n3 * tu2));
  double Szx = fS * (3.0 * rdn * (nu * (tu3 * rd1 + tu1 * rd3) - 5.0 * rd3 * rd1 * rdu) +
                     3.0 * nu * rdu * (n3 * rd1 + n1 * rd3) +
                     c6 * (3.0 * nuv * rd3 * rd1 + n3 * tu1 + n1 * tu3));
  sig_acc[0] += Dxx - Sxx;
  sig_acc[1] += Dyy - Syy;
  sig_acc[2] += Dzz - Szz;
  sig_acc[3] += Dxy - Sxy;
  sig_acc[4] += Dyz - Syz;
  sig_acc[5] += Dzx - Szx;
}

/* stress_mod.f90:511-781 */
void qo_internal_points(int n, int ne, const double* x, const double* y, const double* z,
                        const int* nconec, const int* kconec, const qo_material_t* m,
                        const double* eta1, const double* eta2, const double* eta3, const double* xja,
                        const double* xg1, const double* xg2, const double* xg3, const double* ff,
                        const double* FI, const double* DFI, int l,
                        const double* cx, const double* cy, const double* cz, int variant,
                        double* u_int, double* sigma_int, double* prin, double* vm,
                        long long* counts) {
  (void)n;
  const int NG = 40;
  if (l <= 0) return;
  double* el_c = (double*)malloc(sizeof(double) * 4 * ne);
  int* far_start = (int*)malloc(sizeof(int) * (ne + 1));
  int tot_far = 0;
  for (int e = 0; e < ne; ++e) {
    far_start[e] = tot_far;
    tot_far += (nconec[e] == 9) ? 36 : 16;
  }
  far_start[ne] = tot_far;
  qo_gp_t* far = (qo_gp_t*)malloc(sizeof(qo_gp_t) * (size_t)tot_far);
  qo_gp_t* near = (qo_gp_t*)malloc(sizeof(qo_gp_t) * (size_t)ne * NG * NG);
  double xgl[40], wgl[40];
  qo_gauss_legendre(NG, xgl, wgl);

  for (int e = 0; e < ne; ++e) {
    int nc = nconec[e];
    double xj1[9], xj2[9], xj3[9];
    gather_coords(e, nconec, kconec, x, y, z, xj1, xj2, xj3);
    double ccx = 0.0, ccy = 0.0, ccz = 0.0;
    for (int mm = 0; mm < nc; ++mm) {
      ccx += xj1[mm];
      ccy += xj2[mm];
      ccz += xj3[mm];
    }
    ccx /= (double)nc;
    ccy /= (double)nc;
    ccz /= (double)nc;
    double dmax2 = 0.0;
    for (int mm = 0; mm < nc; ++mm) {
      double dx = xj1[mm] - ccx, dy = xj2[mm] - ccy, dz = xj3[mm] - ccz;
      double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 > dmax2) dmax2 = d2;
    }
    el_c[4 * e + 0] = ccx;
    el_c[4 * e + 1] = ccy;
    el_c[4 * e + 2] = ccz;
    el_c[4 * e + 3] = (3.5 * 3.5) * dmax2;
    int npg = (nc == 9) ? 6 : 4;
    int igp = far_start[e];
    for (int i = 0; i < npg; ++i)
      for (int j = 0; j < npg; ++j, ++igp) {
        int k = npg * i + j;
        double wg = (nc == 9) ? ome_quad[i] * ome_quad[j] * xja[36 * e + k]
                              : omei_tri[i] * omej_tri[j] * xja[36 * e + k];
        qo_gp_t* gp = &far[igp];
        gp->x = xg1[36 * e + k];
        gp->y = xg2[36 * e + k];
        gp->z = xg3[36 * e + k];
        gp->nx = eta1[36 * e + k];
        gp->ny = eta2[36 * e + k];
        gp->nz = eta3[36 * e + k];
        double ug[3] = {0, 0, 0}, tg[3] = {0, 0, 0};
        for (int mm = 0; mm < nc; ++mm) {
          int ik = kconec[e * 9 + mm] - 1;
          double fv = ff[(36 * e + k) * 9 + mm];
          for (int d = 0; d < 3; ++d) {
            ug[d] += fv * FI[3 * ik + d];
            tg[d] += fv * DFI[e * 27 + 3 * mm + d];
          }
        }
        for (int d = 0; d < 3; ++d) {
          gp->u[d] = ug[d] * wg;
          gp->t[d] = tg[d] * wg;
        }
      }
  }

#pragma omp parallel for schedule(static)
  for (int e = 0; e < ne; ++e) {
    int nc = nconec[e];
    double xj1[9], xj2[9], xj3[9];
    gather_coords(e, nconec, kconec, x, y, z, xj1, xj2, xj3);
    for (int i = 0; i < NG; ++i)
      for (int j = 0; j < NG; ++j) {
        qo_gp_t* gp = &near[(size_t)e * NG * NG + i * NG + j];
        double g1, g2, tri_jac;
        if (nc == 9) {
          g1 = xgl[i];
          g2 = xgl[j];
          tri_jac = 1.0;
        } else {
          g1 = 0.5 * (1.0 + xgl[i]);
          g2 = 0.25 * (1.0 - xgl[i]) * (1.0 + xgl[j]);
          tri_jac = 0.125 * (1.0 - xgl[i]);
        }
        double f[9], fd1[9], fd2[9];
        qo_shape(nc, g1, g2, f, fd1, fd2);
        double xf[3] = {0, 0, 0}, dx1[3] = {0, 0, 0}, dx2[3] = {0, 0, 0};
        double ug[3] = {0, 0, 0}, tg[3] = {0, 0, 0};
        for (int mm = 0; mm < nc; ++mm) {
          int ik = kconec[e * 9 + mm] - 1;
          double fv = f[mm];
          xf[0] += fv * xj1[mm];
          xf[1] += fv * xj2[mm];
          xf[2] += fv * xj3[mm];
          dx1[0] += fd1[mm] * xj1[mm];
          dx1[1] += fd1[mm] * xj2[mm];
          dx1[2] += fd1[mm] * xj3[mm];
          dx2[0] += fd2[mm] * xj1[mm];
          dx2[1] += fd2[mm] * xj2[mm];
          dx2[2] += fd2[mm] * xj3[mm];
          for (int d = 0; d < 3; ++d) {
            ug[d] += fv * FI[3 * ik + d];
            tg[d] += fv * DFI[e * 27 + 3 * mm + d];
          }
        }
        double cv[3] = {dx1[1] * dx2[2] - dx1[2] * dx2[1], dx1[2] * dx2[0] - dx1[0] * dx2[2],
                        dx1[0] * dx2[1] - dx1[1] * dx2[0]};
        double jac = sqrt(cv[0] * cv[0] + cv[1] * cv[1] + cv[2] * cv[2]);
        double wg = wgl[i] * wgl[j] * tri_jac * jac;
        gp->x = xf[0];
        gp->y = xf[1];
        gp->z = xf[2];
        gp->nx = cv[0] / jac;
        gp->ny = cv[1] / jac;
        gp->nz = cv[2] / jac;
        for (int d = 0; d < 3; ++d) {
          gp->u[d] = ug[d] * wg;
          gp->t[d] = tg[d] * wg;
        }
      }
  }

  const double nu = m->nu, mu = m->mu;
  const double c4 = 3.0 - 4.0 * nu, c6 = 1.0 - 2.0 * nu, pr4 = 1.0 - 4.0 * nu;
  const double cU = 1.0 / (16.0 * QO_PI * (1.0 - nu) * mu);
  const double cT = -1.0 / (8.0 * QO_PI * (1.0 - nu));
  const double cD = 1.0 / (8.0 * QO_PI * (1.0 - nu));
  const double cS = mu / (4.0 * QO_PI * (1.0 - nu));
  long long nfar = 0, nnear = 0;
#pragma omp parallel for schedule(guided) reduction(+ : nfar, nnear)
  for (int ip = 0; ip < l; ++ip) {
    double u[3] = {0, 0, 0}, s[6] = {0, 0, 0, 0, 0, 0};
    for (int e = 0; e < ne; ++e) {
      double dx = cx[ip] - el_c[4 * e], dy = cy[ip] - el_c[4 * e + 1], dz = cz[ip] - el_c[4 * e + 2];
      double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 > el_c[4 * e + 3]) {
        for (int g = far_start[e]; g < far_start[e + 1]; ++g)
          somigliana_gp(&far[g], cx[ip], cy[ip], cz[ip], nu, c4, c6, pr4, cU, cT, cD, cS, variant, u, s);
        nfar += far_start[e + 1] - far_start[e];
      } else {
        const qo_gp_t* base = &near[(size_t)e * NG * NG];
        for (int g = 0; g < NG * NG; ++g)
          somigliana_gp(&base[g], cx[ip], cy[ip], cz[ip], nu, c4, c6, pr4, cU, cT, cD, cS, variant, u, s);
        nnear += NG * NG;
      }
    }
    for (int d = 0; d < 3; ++d) u_int[3 * ip + d] = u[d];
    for (int d = 0; d < 6; ++d) sigma_int[6 * ip + d] = s[d];
    qo_principal_vm(s, prin + 3 * ip, vm + ip);
  }
  if (counts) {
    counts[0] = nfar;
    counts[1] = nnear;
  }
  free(el_c);
  free(far_start);
  free(far);
  free(near);
}

/* stress_mod.f90:119-170; L(i,j) at L[i*3+j] (row i = local axis) */
static void axes_cos(int ncone, const double fd1[9], const double fd2[9], const double co[27],
                     double L[9], double CL[4], double* v1o, double* v2o) {
  double g1[3] = {0, 0, 0}, g2[3] = {0, 0, 0}, g3[3];
  for (int a = 0; a < ncone; ++a)
    for (int i = 0; i < 3; ++i) {
      g1[i] += fd1[a] * co[3 * a + i];
      g2[i] += fd2[a] * co[3 * a + i];
    }
  double v1 = sqrt(g1[0] * g1[0] + g1[1] * g1[1] + g1[2] * g1[2]);
  double v2 = sqrt(g2[0] * g2[0] + g2[1] * g2[1] + g2[2] * g2[2]);
  g3[0] = g1[1] * g2[2] - g1[2] * g2[1];
  g3[1] = g1[2] * g2[0] - g1[0] * g2[2];
  g3[2] = g1[0] * g2[1] - g1[1] * g2[0];
  double jcb = sqrt(g3[0] * g3[0] + g3[1] * g3[1] + g3[2] * g3[2]);
  for (int i = 0; i < 3; ++i) {
    L[6 + i] = g3[i] / jcb;
    L[0 + i] = g1[i] / v1;
  }
  L[3 + 0] = L[6 + 1] * L[0 + 2] - L[6 + 2] * L[0 + 1];
  L[3 + 1] = L[6 + 2] * L[0 + 0] - L[6 + 0] * L[0 + 2];
  L[3 + 2] = L[6 + 0] * L[0 + 1] - L[6 + 1] * L[0 + 0];
  double cth = (g1[0] * g2[0] + g1[1] * g2[1] + g1[2] * g2[2]) / (v1 * v2);
  if (cth > 1.0) cth = 1.0;
  if (cth < -1.0) cth = -1.0;
  double sth = sqrt(fmax(0.0, 1.0 - cth * cth));
  if (sth < 1.0e-12) sth = 1.0e-12;
  CL[0] = 1.0 / v1;              /* CL(1,1) */
  CL[1] = -cth / (v1 * sth);     /* CL(1,2) */
  CL[2] = 0.0;                   /* CL(2,1) */
  CL[3] = 1.0 / (v2 * sth);      /* CL(2,2) */
  *v1o = v1;
  *v2o = v2;
}

/* stress_mod.f90:179-292, 366-406 */
void qo_boundary_stresses(int n, int ne, const double* x, const double* y, const double* z,
                          const int* nconec, const int* kconec, const int* nesim, const int* ke,
                          const qo_material_t* m, const double* FI, const double* DFI,
                          double* sigma_elem, double* sigma_node, double* prin, double* vm) {
  const double rge = m->mu, xnu = m->nu;
#pragma omp parallel for schedule(dynamic)
  for (int ie = 0; ie < ne; ++ie) {
    int nc = nconec[ie];
    double co[27], ue[27], te[27];
    for (int a = 0; a < nc; ++a) {
      int ik = kconec[ie * 9 + a] - 1;
      co[3 * a] = x[ik];
      co[3 * a + 1] = y[ik];
      co[3 * a + 2] = z[ik];
      for (int d = 0; d < 3; ++d) {
        ue[3 * a + d] = FI[3 * ik + d];
        te[3 * a + d] = DFI[ie * 27 + 3 * a + d];
      }
    }
    static const double q1[9] = {-1.0, 0.0, 1.0, 1.0, 1.0, 0.0, -1.0, -1.0, 0.0};
    static const double q2[9] = {-1.0, -1.0, -1.0, 0.0, 1.0, 1.0, 1.0, 0.0, 0.0};
    static const double t1[6] = {1.0, 0.0, 0.0, 0.5, 0.0, 0.5};
    static const double t2[6] = {0.0, 1.0, 0.0, 0.5, 0.5, 0.0};
    const double c_eps = 2.0 * rge / (1.0 - xnu);
    const double c_nu = 2.0 * rge * xnu / (1.0 - xnu);
    const double c_sig3 = xnu / (1.0 - xnu);
    double* se = sigma_elem + (size_t)ie * 54;
    for (int i = 0; i < 54; ++i) se[i] = 0.0;
    for (int k = 0; k < nc; ++k) {
      double g1 = (nc == 9) ? q1[k] : t1[k];
      double g2 = (nc == 9) ? q2[k] : t2[k];
      double f[9], fd1[9], fd2[9], L[9], CL[4], v1, v2;
      qo_shape(nc, g1, g2, f, fd1, fd2);
      axes_cos(nc, fd1, fd2, co, L, CL, &v1, &v2);
      double dn1[9], dn2[9], ul[27];
      for (int a = 0; a < nc; ++a) {
        dn1[a] = CL[0] * fd1[a] + CL[2] * fd2[a];
        dn2[a] = CL[1] * fd1[a] + CL[3] * fd2[a];
      }
      for (int a = 0; a < nc; ++a)
        for (int i = 0; i < 3; ++i)
          ul[3 * a + i] = L[3 * i] * ue[3 * a] + L[3 * i + 1] * ue[3 * a + 1] + L[3 * i + 2] * ue[3 * a + 2];
      double u11 = 0, u22 = 0, u12 = 0, u21 = 0;
      for (int a = 0; a < nc; ++a) {
        u11 += dn1[a] * ul[3 * a];
        u22 += dn2[a] * ul[3 * a + 1];
        u12 += dn2[a] * ul[3 * a];
        u21 += dn1[a] * ul[3 * a + 1];
      }
      double e11 = u11, e22 = u22, e12 = 0.5 * (u12 + u21);
      double tl[3];
      for (int i = 0; i < 3; ++i)
        tl[i] = L[3 * i] * te[3 * k] + L[3 * i + 1] * te[3 * k + 1] + L[3 * i + 2] * te[3 * k + 2];
      double S[9]; /* S(i,j) at S[i*3+j] */
      S[0 * 3 + 2] = tl[0];
      S[2 * 3 + 0] = tl[0];
      S[1 * 3 + 2] = tl[1];
      S[2 * 3 + 1] = tl[1];
      S[2 * 3 + 2] = tl[2];
      S[0] = c_eps * e11 + c_nu * e22 + c_sig3 * S[8];
      S[4] = c_eps * e22 + c_nu * e11 + c_sig3 * S[8];
      S[1] = 2.0 * rge * e12;
      S[3] = S[1];
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
  for (int in = 0; in < n; ++in) {
    double* sn = sigma_node + 6 * in;
    for (int d = 0; d < 6; ++d) sn[d] = 0.0;
    int ns = nesim[in] < 8 ? nesim[in] : 8;  /* stress_mod.f90:385: min(nesim, 8) */
    if (ns > 0) {
      for (int j = 0; j < ns; ++j) {
        int ie = ke[(long)in * g_kecap + j] - 1;
        for (int k = 0; k < nconec[ie]; ++k)
          if (kconec[ie * 9 + k] - 1 == in) {
            for (int d = 0; d < 6; ++d) sn[d] += sigma_elem[(size_t)ie * 54 + 6 * k + d];
            break;
          }
      }
      for (int d = 0; d < 6; ++d) sn[d] /= (double)ns;
    }
    qo_principal_vm(sn, prin + 3 * in, vm + in);
  }
}
