// qd_lapack.cpp -- the LAPACK-ABI DGESV that solver_mod.f90:28 calls, backed
// by the device LU of qd_lu.hip.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <mutex>

#include "qd_internal.h"

namespace {
std::mutex g_mu;
qd_handle g_h = nullptr;

qd_handle lapack_handle() {
  if (g_h) return g_h;
  int dev = 0;
  if (const char* e = std::getenv("QUADESTH_DEVICE")) dev = std::atoi(e);
  if (qd_create(&g_h, dev) != QD_OK) g_h = nullptr;
  return g_h;
}
}  // namespace

extern "C" {

int qd_dgesv_device(qd_handle h, int n, int nrhs, double* a_dev, long long lda, int* ipiv, double* b_dev,
                    long long ldb) {
  if (!h) return QD_ERR_ARG;
  if (n < 0) return qd::set_error(h, -1, "dgesv: n=%d", n);
  if (nrhs < 0) return qd::set_error(h, -2, "dgesv: nrhs=%d", nrhs);
  if (lda < (n > 1 ? n : 1)) return qd::set_error(h, -4, "dgesv: lda=%lld", lda);
  if (ldb < (n > 1 ? n : 1)) return qd::set_error(h, -7, "dgesv: ldb=%lld", ldb);
  if (n == 0) return QD_OK;
  if (!a_dev || !ipiv || (nrhs > 0 && !b_dev)) return QD_ERR_ARG;
  QD_HIP(h, hipSetDevice(h->device));
  return qd::dgesv_device(h, n, nrhs, a_dev, lda, ipiv, b_dev, ldb);
}

// LAPACK DGESV: info = 0 success, -i bad argument i, +i zero pivot U(i,i).
void dgesv_(const int* n, const int* nrhs, double* a, const int* lda, int* ipiv, double* b, const int* ldb,
            int* info) {
  std::lock_guard<std::mutex> lock(g_mu);
  *info = 0;
  if (*n < 0) *info = -1;
  else if (*nrhs < 0) *info = -2;
  else if (*lda < (*n > 1 ? *n : 1)) *info = -4;
  else if (*ldb < (*n > 1 ? *n : 1)) *info = -7;
  if (*info != 0 || *n == 0) return;
  qd_handle h = lapack_handle();
  if (!h) {
    *info = -100;
    return;
  }
  const size_t N = (size_t)*n;
  qd::DevBuf<double> bufa, bufb;
  if (hipSetDevice(h->device) != hipSuccess || bufa.ensure(N * N) != hipSuccess ||
      bufb.ensure(N * (size_t)(*nrhs > 0 ? *nrhs : 1)) != hipSuccess) {
    *info = -100;
    return;
  }
  double* da = bufa.p;
  double* db = bufb.p;
  bool ok = hipMemcpy2D(da, N * sizeof(double), a, (size_t)*lda * sizeof(double), N * sizeof(double), N,
                        hipMemcpyHostToDevice) == hipSuccess;
  if (ok && *nrhs > 0)
    ok = hipMemcpy2D(db, N * sizeof(double), b, (size_t)*ldb * sizeof(double), N * sizeof(double), *nrhs,
                     hipMemcpyHostToDevice) == hipSuccess;
  int rc = ok ? qd_dgesv_device(h, *n, *nrhs, da, (long long)N, ipiv, db, (long long)N) : QD_ERR_HIP;
  if (rc >= 0) {
    ok = hipMemcpy2D(a, (size_t)*lda * sizeof(double), da, N * sizeof(double), N * sizeof(double), N,
                     hipMemcpyDeviceToHost) == hipSuccess;
    if (ok && rc == 0 && *nrhs > 0)
      ok = hipMemcpy2D(b, (size_t)*ldb * sizeof(double), db, N * sizeof(double), N * sizeof(double), *nrhs,
                       hipMemcpyDeviceToHost) == hipSuccess;
    *info = ok ? rc : -100;
  } else {
    *info = -100;
  }
}

}  // extern "C"
