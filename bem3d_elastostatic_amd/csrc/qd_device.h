// qd_device.h -- device helpers shared by the kernels (header-only).
#pragma once

#include <hip/hip_runtime.h>

namespace qd {

// Shape functions and intrinsic derivatives, assembly_mod.f90:247-303.
static __device__ __forceinline__ void shape(int ncone, double g1, double g2, double* ff, double* fd1,
                                      double* fd2) {
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
    for (int i = 6; i < 9; ++i) ff[i] = fd1[i] = fd2[i] = 0.0;
  }
}

// Unnormalised normal in the reference's dc/dct form (assembly_mod.f90:311-329).
static __device__ __forceinline__ void normal_dc(int ncone, const double* fd1, const double* fd2,
                                          const double* xj1, const double* xj2, const double* xj3,
                                          double& s1, double& s2, double& s3) {
  double a = 0.0, b = 0.0, c = 0.0;
  for (int jk = 0; jk < ncone; ++jk) {
    double aux = 0.0, bux = 0.0, cux = 0.0;
    for (int ik = 0; ik < ncone; ++ik) {
      double d = fd1[ik] * fd2[jk] - fd2[ik] * fd1[jk];
      aux += xj2[ik] * d;
      bux += xj3[ik] * d;
      cux += xj1[ik] * d;
    }
    a += aux * xj3[jk];
    b += bux * xj1[jk];
    c += cux * xj2[jk];
  }
  s1 = a;
  s2 = b;
  s3 = c;
}

// 1/r to full FP64 accuracy: hardware reciprocal square root refined by two
// Newton steps.
static __device__ __forceinline__ double inv_sqrt(double r2) {
  double y = __builtin_amdgcn_rsq(r2);
  y = y * (1.5 - 0.5 * r2 * y * y);
  y = y * (1.5 - 0.5 * r2 * y * y);
  return y;
}

}  // namespace qd
