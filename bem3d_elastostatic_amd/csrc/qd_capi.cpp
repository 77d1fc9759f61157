// qd_capi.cpp -- host side of the libquadesth C ABI: argument checks, the
// host-side preprocessing of the mesh and boundary conditions, and the
// dispatch into the device phases.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <set>

#include "qd_internal.h"

namespace qd {

int set_error(qd_context* h, int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (h) h->err = buf;
  return code;
}

// Sub-triangles of the singular Duffy split, assembly_mod.f90:399-489.
static void build_subtri(std::vector<SubTri>& tab) {
  tab.assign(2 * 10 * 4, SubTri{});
  // quad9 (kind 0)
  const double gi1[9] = {-1.0, 0.0, 1.0, 1.0, 1.0, 0.0, -1.0, -1.0, 0.0};
  const double gi2[9] = {-1.0, -1.0, -1.0, 0.0, 1.0, 1.0, 1.0, 0.0, 0.0};
  double gii1[16], gii2[16];
  for (int i = 1; i <= 8; ++i) {
    gii1[i] = gi1[i - 1];
    gii2[i] = gi2[i - 1];
  }
  for (int i = 9; i <= 15; ++i) {
    gii1[i] = gi1[i - 9];
    gii2[i] = gi2[i - 9];
  }
  for (int nodo = 1; nodo <= 9; ++nodo) {
    SubTri* t = &tab[(0 * 10 + nodo) * 4];
    double v10 = gi1[nodo - 1], v20 = gi2[nodo - 1];
    auto put = [&](int s, double a1, double b1, double a2, double b2, double area) {
      t[s].v1[0] = v10;
      t[s].v2[0] = v20;
      t[s].v1[1] = a1;
      t[s].v2[1] = b1;
      t[s].v1[2] = a2;
      t[s].v2[2] = b2;
      t[s].area = area;
    };
    if (nodo == 1 || nodo == 3 || nodo == 5 || nodo == 7) {
      for (int k = 1; k <= 2; ++k) {
        int m = 2 * k;
        put(k - 1, gii1[nodo + m], gii2[nodo + m], gii1[nodo + m + 2], gii2[nodo + m + 2], 2.0);
      }
    } else if (nodo == 9) {
      for (int k = 1; k <= 4; ++k) {
        int m = 2 * k - 1;
        put(k - 1, gii1[m], gii2[m], gii1[m + 2], gii2[m + 2], 1.0);
      }
    } else {
      for (int k = 1; k <= 3; ++k) {
        int m = 2 * k - 1;
        double a1 = gii1[nodo + m], b1 = gii2[nodo + m];
        double a2 = gii1[nodo + m + 2], b2 = gii2[nodo + m + 2];
        double area = 0.5 * (a1 * b2 + v10 * b1 + a2 * v20 - a1 * v20 - a2 * b1 - v10 * b2);
        put(k - 1, a1, b1, a2, b2, area);
      }
    }
  }
  // tri6 (kind 1)
  const double t1[6] = {1.0, 0.0, 0.0, 0.5, 0.0, 0.5};
  const double t2[6] = {0.0, 1.0, 0.0, 0.5, 0.5, 0.0};
  const int sub[7][2][2] = {{{0, 0}, {0, 0}}, {{2, 3}, {0, 0}}, {{3, 1}, {0, 0}}, {{1, 2}, {0, 0}},
                            {{2, 3}, {3, 1}}, {{3, 1}, {1, 2}}, {{1, 2}, {2, 3}}};
  for (int nodo = 1; nodo <= 6; ++nodo) {
    SubTri* t = &tab[(1 * 10 + nodo) * 4];
    int nsub = nodo <= 3 ? 1 : 2;
    double area = nodo <= 3 ? 0.5 : 0.25;
    for (int s = 0; s < nsub; ++s) {
      t[s].v1[0] = t1[nodo - 1];
      t[s].v2[0] = t2[nodo - 1];
      t[s].v1[1] = t1[sub[nodo][s][0] - 1];
      t[s].v2[1] = t2[sub[nodo][s][0] - 1];
      t[s].v1[2] = t1[sub[nodo][s][1] - 1];
      t[s].v2[2] = t2[sub[nodo][s][1] - 1];
      t[s].area = area;
    }
  }
}

static int nsub_of(int nconec, int nodo) {
  if (nconec == 9) {
    if (nodo == 9) return 4;
    return (nodo % 2 == 1) ? 2 : 3;
  }
  return nodo <= 3 ? 1 : 2;
}

static int upload_i(qd_context* h, DevBuf<int>& d, const std::vector<int>& v) {
  QD_ALLOC(h, d, std::max<size_t>(v.size(), 1));
  if (!v.empty()) QD_HIP(h, hipMemcpy(d.p, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
  return QD_OK;
}
static int upload_d(qd_context* h, DevBuf<double>& d, const double* v, size_t cnt) {
  QD_ALLOC(h, d, std::max<size_t>(cnt, 1));
  if (cnt) QD_HIP(h, hipMemcpy(d.p, v, cnt * sizeof(double), hipMemcpyHostToDevice));
  return QD_OK;
}

}  // namespace qd

using namespace qd;

extern "C" {

// Message of the last qd_create failure on this thread (qd_last_error(NULL)).
static thread_local std::string g_create_err = "no handle";

static int create_failed(int code, const std::string& msg) {
  g_create_err = msg;
  return code;
}

int qd_create(qd_handle* out, int device) {
  if (!out) return create_failed(QD_ERR_ARG, "qd_create: out is NULL");
  *out = nullptr;
  int ndev = 0;
  const hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0)
    return create_failed(QD_ERR_HIP, std::string("qd_create: no HIP device (") +
                                         (e != hipSuccess ? hipGetErrorString(e) : "0 devices") + ")");
  if (device < 0 || device >= ndev)
    return create_failed(QD_ERR_ARG, "qd_create: device " + std::to_string(device) + " out of range [0, " +
                                         std::to_string(ndev) + ")");
  if (hipSetDevice(device) != hipSuccess) return create_failed(QD_ERR_HIP, "qd_create: hipSetDevice failed");
  qd_context* h = new qd_context();
  h->device = device;
  std::string serr;
  if (noalloc_init(device, &serr) != QD_OK) {
    delete h;
    return create_failed(QD_ERR_HIP, "qd_create: " + serr);
  }
  std::vector<SubTri> st;
  build_subtri(st);
  if (h->subtri.ensure(st.size()) != hipSuccess ||
      hipMemcpy(h->subtri.p, st.data(), st.size() * sizeof(SubTri), hipMemcpyHostToDevice) !=
          hipSuccess) {
    delete h;
    return create_failed(QD_ERR_HIP, "qd_create: device allocation failed");
  }
  *out = h;
  return QD_OK;
}

void qd_destroy(qd_handle h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  (void)hipDeviceSynchronize();
  lu_free(h);
  delete h;
}

const char* qd_last_error(qd_handle h) { return h ? h->err.c_str() : g_create_err.c_str(); }

int qd_set_stream(qd_handle h, void* s) {
  if (!h) return QD_ERR_ARG;
  h->stream = (hipStream_t)s;
  return QD_OK;
}

int qd_synchronize(qd_handle h) {
  if (!h) return QD_ERR_ARG;
  QD_HIP(h, hipSetDevice(h->device));
  QD_HIP(h, hipStreamSynchronize(h->stream));
  return QD_OK;
}

int qd_set_mesh(qd_handle h, int n, int ne, const double* x, const double* y, const double* z,
                const int* nconec, const int* kconec) {
  if (!h) return QD_ERR_ARG;
  if (n <= 0 || ne <= 0 || !x || !y || !z || !nconec || !kconec)
    return set_error(h, QD_ERR_ARG, "qd_set_mesh: n=%d ne=%d or null array", n, ne);
  if ((long long)3 * n > 2147483647LL) return set_error(h, QD_ERR_ARG, "qd_set_mesh: 3n overflows int");
  QD_HIP(h, hipSetDevice(h->device));
  for (int e = 0; e < ne; ++e) {
    if (nconec[e] != 9 && nconec[e] != 6)
      return set_error(h, QD_ERR_MESH, "element %d: nconec=%d (expected 9 or 6)", e + 1, nconec[e]);
    for (int k = 0; k < nconec[e]; ++k) {
      int v = kconec[e * 9 + k];
      if (v < 1 || v > n)
        return set_error(h, QD_ERR_MESH, "element %d node %d: %d out of 1..%d", e + 1, k + 1, v, n);
    }
  }
  h->n = n;
  h->ne = ne;
  h->hx.assign(x, x + n);
  h->hy.assign(y, y + n);
  h->hz.assign(z, z + n);
  h->hnconec.assign(nconec, nconec + ne);
  h->hkconec.assign(kconec, kconec + 9 * (size_t)ne);
  for (int e = 0; e < ne; ++e)
    for (int k = h->hnconec[e]; k < 9; ++k) h->hkconec[e * 9 + k] = 0;
  // nesim and the elements of each node in ascending order (io_mod.f90:124-135).
  // The reference keeps the first 8 in ke(8, n) and its kode0 search and
  // reorganize_solution then read past ke for a node in more than 8
  // elements (assembly_mod.f90:150-151, io_mod.f90:170-171); the list here
  // holds every element, so those searches see them all.
  h->nesim.assign(n, 0);
  for (int e = 0; e < ne; ++e)
    for (int j = 0; j < h->hnconec[e]; ++j) h->nesim[h->hkconec[e * 9 + j] - 1] += 1;
  h->ke_ptr.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    if (h->nesim[i] == 0) return set_error(h, QD_ERR_MESH, "node %d belongs to no element", i + 1);
    h->ke_ptr[i + 1] = h->ke_ptr[i] + h->nesim[i];
  }
  h->ke.assign((size_t)h->ke_ptr[n], 0);
  {
    std::vector<int> fill(h->ke_ptr.begin(), h->ke_ptr.end() - 1);
    for (int e = 0; e < ne; ++e)
      for (int j = 0; j < h->hnconec[e]; ++j) h->ke[fill[h->hkconec[e * 9 + j] - 1]++] = e + 1;
  }
  // singular pairs, node-major, elements ascending (the ke order)
  h->h_pairs.clear();
  h->h_pair_ptr.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    h->h_pair_ptr[i] = (int)h->h_pairs.size();
    for (int j = 0; j < h->nesim[i]; ++j) {
      int e = h->ke[h->ke_ptr[i] + j] - 1;
      int nodo = 0;
      for (int k = 0; k < h->hnconec[e]; ++k)
        if (h->hkconec[e * 9 + k] - 1 == i) {
          nodo = k + 1;
          break;
        }
      h->h_pairs.push_back(SingPair{i, e, nodo, nsub_of(h->hnconec[e], nodo)});
    }
  }
  h->h_pair_ptr[n] = (int)h->h_pairs.size();
  build_schedule(h);

  int rc;
  if ((rc = upload_d(h, h->x, x, n))) return rc;
  if ((rc = upload_d(h, h->y, y, n))) return rc;
  if ((rc = upload_d(h, h->z, z, n))) return rc;
  if ((rc = upload_i(h, h->nconec, h->hnconec))) return rc;
  std::vector<int> k0(h->hkconec.size());
  for (size_t i = 0; i < k0.size(); ++i) k0[i] = h->hkconec[i] - 1;
  if ((rc = upload_i(h, h->kconec0, k0))) return rc;
  if ((rc = upload_i(h, h->chunk_start, h->h_chunk_start))) return rc;
  if ((rc = upload_i(h, h->colour_chunks, h->h_colour_chunks))) return rc;
  if ((rc = upload_i(h, h->pair_ptr, h->h_pair_ptr))) return rc;
  QD_ALLOC(h, h->pairs, h->h_pairs.size());
  QD_HIP(h, hipMemcpy(h->pairs.p, h->h_pairs.data(), h->h_pairs.size() * sizeof(SingPair),
                      hipMemcpyHostToDevice));
  h->have_mesh = true;
  h->have_bc = false;
  h->have_geom = false;
  h->have_sys = false;
  lu_invalidate(h);
  return QD_OK;
}

int qd_set_coordinates(qd_handle h, const double* x, const double* y, const double* z) {
  if (!h || !x || !y || !z) return QD_ERR_ARG;
  if (!h->have_mesh) return set_error(h, QD_ERR_STATE, "qd_set_coordinates before qd_set_mesh");
  QD_HIP(h, hipSetDevice(h->device));
  const int n = h->n;
  h->hx.assign(x, x + n);
  h->hy.assign(y, y + n);
  h->hz.assign(z, z + n);
  int rc;
  if ((rc = upload_d(h, h->x, x, n))) return rc;
  if ((rc = upload_d(h, h->y, y, n))) return rc;
  if ((rc = upload_d(h, h->z, z, n))) return rc;
  h->have_geom = false;
  h->have_sys = false;
  lu_invalidate(h);
  return QD_OK;
}

int qd_set_material(qd_handle h, double mu, double nu) {
  if (!h) return QD_ERR_ARG;
  if (!(mu > 0.0) || !(nu < 1.0) || !(nu > -1.0))
    return set_error(h, QD_ERR_ARG, "qd_set_material: mu=%g nu=%g", mu, nu);
  // data_mod.f90:97-108
  const double pi = 3.14159265358979323846;
  h->mat.mu = mu;
  h->mat.nu = nu;
  h->mat.young = 2.0 * mu * (1.0 + nu);
  h->mat.c3 = 1.0 / (16.0 * pi * (1.0 - nu) * mu);
  h->mat.c4 = 3.0 - 4.0 * nu;
  h->mat.c5 = -1.0 / (8.0 * pi * (1.0 - nu));
  h->mat.c6 = 1.0 - 2.0 * nu;
  h->have_mat = true;
  h->have_sys = false;
  lu_invalidate(h);
  return QD_OK;
}

int qd_set_bc(qd_handle h, const int* kode, const double* bc) {
  if (!h || !kode || !bc) return QD_ERR_ARG;
  if (!h->have_mesh) return set_error(h, QD_ERR_STATE, "qd_set_bc before qd_set_mesh");
  QD_HIP(h, hipSetDevice(h->device));
  std::vector<ElemInfo> info;
  std::vector<double> evp, nvp;
  std::vector<int> nk0;
  int rc = bc_tables(h, kode, bc, info, evp, nk0, nvp);
  if (rc) return rc;
  h->hkode.assign(kode, kode + 3 * (size_t)h->ne);
  h->hbc.assign(bc, bc + 27 * (size_t)h->ne);
  build_gschedule(h, info);
  std::vector<unsigned> ops;
  build_write_plan(h, info, ops);
  QD_ALLOC(h, h->elop, ops.size());
  QD_HIP(h, hipMemcpy(h->elop.p, ops.data(), ops.size() * sizeof(unsigned), hipMemcpyHostToDevice));
  if ((rc = upload_i(h, h->unwritten, h->h_unwritten))) return rc;
  QD_ALLOC(h, h->einfo, info.size());
  QD_HIP(h, hipMemcpy(h->einfo.p, info.data(), info.size() * sizeof(ElemInfo), hipMemcpyHostToDevice));
  if ((rc = upload_d(h, h->ebc, bc, 27 * (size_t)h->ne))) return rc;
  if ((rc = upload_d(h, h->evp, evp.data(), evp.size()))) return rc;
  if ((rc = upload_i(h, h->node_k0, nk0))) return rc;
  if ((rc = upload_d(h, h->node_vp, nvp.data(), nvp.size()))) return rc;
  if ((rc = upload_i(h, h->gchunks, h->h_gchunks))) return rc;
  h->have_bc = true;
  h->have_sys = false;
  lu_invalidate(h);
  return QD_OK;
}

int qd_precompute_element_geometry(qd_handle h, double* eta1, double* eta2, double* eta3,
                                   double* xja, double* xg1, double* xg2, double* xg3, double* ff) {
  if (!h) return QD_ERR_ARG;
  if (!h->have_mesh) return set_error(h, QD_ERR_STATE, "geometry before qd_set_mesh");
  QD_HIP(h, hipSetDevice(h->device));
  int rc = geometry_run(h);
  if (rc) return rc;
  h->have_geom = true;
  const size_t blk = 36 * (size_t)h->ne;
  double* outs[7] = {eta1, eta2, eta3, xja, xg1, xg2, xg3};
  bool any = ff != nullptr;
  for (double* o : outs) any |= o != nullptr;
  if (any) {
    QD_HIP(h, hipStreamSynchronize(h->stream));
    for (int i = 0; i < 7; ++i)
      if (outs[i])
        QD_HIP(h, hipMemcpy(outs[i], h->geo.p + i * blk, blk * sizeof(double), hipMemcpyDeviceToHost));
    if (ff) QD_HIP(h, hipMemcpy(ff, h->geo.p + 7 * blk, 9 * blk * sizeof(double), hipMemcpyDeviceToHost));
  }
  return QD_OK;
}

int qd_ghmat8(qd_handle h) {
  if (!h) return QD_ERR_ARG;
  if (!h->have_mesh || !h->have_mat || !h->have_bc)
    return set_error(h, QD_ERR_STATE, "qd_ghmat8 needs mesh, material and boundary conditions");
  QD_HIP(h, hipSetDevice(h->device));
  if (!h->have_geom) {
    int rc = geometry_run(h);
    if (rc) return rc;
    h->have_geom = true;
  }
  lu_invalidate(h);
  int rc = assembly_run(h, nullptr, true, h->asm_mode == QD_ASSEMBLY_FAR_FP32);
  if (rc) return rc;
  h->have_pre = false;
  if (h->asm_mode == QD_ASSEMBLY_FAR_FP32_FACTOR) {
    // the matrix the mixed-precision factorisations factor: far field in FP32
    if ((rc = assembly_run(h, nullptr, true, true, true))) return rc;
    h->have_pre = true;
  } else {
    h->A_pre.release();
    h->FI_pre.release();
  }
  h->have_sys = true;
  return QD_OK;
}

int qd_set_assembly_precision(qd_handle h, int mode) {
  if (!h) return QD_ERR_ARG;
  if (mode != QD_ASSEMBLY_FP64 && mode != QD_ASSEMBLY_FAR_FP32 && mode != QD_ASSEMBLY_FAR_FP32_FACTOR)
    return set_error(h, QD_ERR_ARG, "qd_set_assembly_precision: mode %d", mode);
  h->asm_mode = mode;
  h->have_sys = false;
  lu_invalidate(h);
  return QD_OK;
}

int qd_ghmat8_rhs(qd_handle h, const double* bc, double* fi) {
  if (!h || !bc || !fi) return QD_ERR_ARG;
  if (!h->have_bc || !h->have_mat) return set_error(h, QD_ERR_STATE, "qd_ghmat8_rhs before qd_set_bc");
  QD_HIP(h, hipSetDevice(h->device));
  if (!h->have_geom) {
    int rc = geometry_run(h);
    if (rc) return rc;
    h->have_geom = true;
  }
  // The same kode with other prescribed values: vp changes with bc, the
  // kode0 flags do not.
  std::vector<ElemInfo> info;
  std::vector<double> evp, nvp;
  std::vector<int> nk0;
  int rc = bc_tables(h, h->hkode.data(), bc, info, evp, nk0, nvp);
  if (rc) return rc;
  DevBuf<double> dbc, devp, dnvp, dfi;
  if ((rc = upload_d(h, dbc, bc, 27 * (size_t)h->ne))) return rc;
  if ((rc = upload_d(h, devp, evp.data(), evp.size()))) return rc;
  if ((rc = upload_d(h, dnvp, nvp.data(), nvp.size()))) return rc;
  QD_ALLOC(h, dfi, 3 * (size_t)h->n);
  std::swap(h->ebc.p, dbc.p);
  std::swap(h->evp.p, devp.p);
  std::swap(h->node_vp.p, dnvp.p);
  // element info flags (needG) may differ: swap too
  DevBuf<ElemInfo> dinfo;
  QD_ALLOC(h, dinfo, info.size());
  QD_HIP(h, hipMemcpy(dinfo.p, info.data(), info.size() * sizeof(ElemInfo), hipMemcpyHostToDevice));
  std::swap(h->einfo.p, dinfo.p);
  std::vector<int> save_g = h->h_gchunks, save_gs = h->h_gcolour_start;
  build_gschedule(h, info);
  DevBuf<int> dg;
  rc = upload_i(h, dg, h->h_gchunks);
  if (!rc) {
    std::swap(h->gchunks.p, dg.p);
    rc = assembly_run(h, dfi.p, false, h->asm_mode == QD_ASSEMBLY_FAR_FP32);
    std::swap(h->gchunks.p, dg.p);
  }
  std::swap(h->ebc.p, dbc.p);
  std::swap(h->evp.p, devp.p);
  std::swap(h->node_vp.p, dnvp.p);
  std::swap(h->einfo.p, dinfo.p);
  h->h_gchunks = save_g;
  h->h_gcolour_start = save_gs;
  if (rc) return rc;
  QD_HIP(h, hipStreamSynchronize(h->stream));
  QD_HIP(h, hipMemcpy(fi, dfi.p, 3 * (size_t)h->n * sizeof(double), hipMemcpyDeviceToHost));
  return QD_OK;
}

int qd_get_system(qd_handle h, double* A, int lda, double* FI) {
  if (!h) return QD_ERR_ARG;
  if (!h->have_sys) return set_error(h, QD_ERR_STATE, "qd_get_system before qd_ghmat8");
  const long long N = 3LL * h->n;
  if (A && lda < N) return set_error(h, QD_ERR_ARG, "lda=%d < N=%lld", lda, N);
  QD_HIP(h, hipSetDevice(h->device));
  QD_HIP(h, hipStreamSynchronize(h->stream));
  if (A)
    QD_HIP(h, hipMemcpy2D(A, (size_t)lda * sizeof(double), h->A.p, (size_t)h->lda * sizeof(double),
                          (size_t)N * sizeof(double), (size_t)N, hipMemcpyDeviceToHost));
  if (FI) QD_HIP(h, hipMemcpy(FI, h->FI.p, (size_t)N * sizeof(double), hipMemcpyDeviceToHost));
  return QD_OK;
}

int qd_system_device(qd_handle h, double** A_dev, long long* lda, double** FI_dev) {
  if (!h) return QD_ERR_ARG;
  if (!h->have_sys) return set_error(h, QD_ERR_STATE, "qd_system_device before qd_ghmat8");
  if (A_dev) *A_dev = h->A.p;
  if (lda) *lda = h->lda;
  if (FI_dev) *FI_dev = h->FI.p;
  return QD_OK;
}

int qd_solve_device(qd_handle h, int precision, int nrhs, double* b_dev, long long ldb, double tol,
                    int max_iter, qd_solve_report* rep) {
  if (!h) return QD_ERR_ARG;
  if (!h->have_sys) return set_error(h, QD_ERR_STATE, "qd_solve before qd_ghmat8");
  if (precision < QD_FP64 || precision > QD_FP16_IR)
    return set_error(h, QD_ERR_ARG, "precision %d", precision);
  if (nrhs < 1 || !b_dev || ldb < 3LL * h->n)
    return set_error(h, QD_ERR_ARG, "nrhs=%d ldb=%lld", nrhs, ldb);
  QD_HIP(h, hipSetDevice(h->device));
  return lu_solve(h, precision, nrhs, b_dev, ldb, tol, max_iter, false, rep);
}

int qd_solve_device_noalloc(qd_handle h, int precision, int nrhs, double* b_dev, long long ldb, double tol,
                            int max_iter, qd_solve_report* rep) {
  if (!h) return QD_ERR_ARG;
  if (!h->have_sys) return set_error(h, QD_ERR_STATE, "qd_solve_device_noalloc before qd_ghmat8");
  if (precision < QD_FP64 || precision > QD_FP16_IR)
    return set_error(h, QD_ERR_ARG, "precision %d", precision);
  if (nrhs < 1 || !b_dev || ldb < 3LL * h->n)
    return set_error(h, QD_ERR_ARG, "nrhs=%d ldb=%lld", nrhs, ldb);
  QD_HIP(h, hipSetDevice(h->device));
  return lu_solve_noalloc(h, precision, nrhs, b_dev, ldb, tol, max_iter, rep);
}

int qd_solve_transpose_device(qd_handle h, int precision, int nrhs, double* c_dev, long long ldc,
                              double tol, int max_iter, qd_solve_report* rep) {
  if (!h) return QD_ERR_ARG;
  if (!h->have_sys) return set_error(h, QD_ERR_STATE, "qd_solve_transpose before qd_ghmat8");
  if (precision < QD_FP64 || precision > QD_FP16_IR)
    return set_error(h, QD_ERR_ARG, "precision %d", precision);
  if (nrhs < 1 || !c_dev || ldc < 3LL * h->n)
    return set_error(h, QD_ERR_ARG, "nrhs=%d ldc=%lld", nrhs, ldc);
  QD_HIP(h, hipSetDevice(h->device));
  return lu_solve(h, precision, nrhs, c_dev, ldc, tol, max_iter, true, rep);
}

int qd_solve(qd_handle h, int precision, int nrhs, double* b, int ldb, double tol, int max_iter,
             qd_solve_report* rep) {
  if (!h) return QD_ERR_ARG;
  if (!h->have_sys) return set_error(h, QD_ERR_STATE, "qd_solve before qd_ghmat8");
  const long long N = 3LL * h->n;
  QD_HIP(h, hipSetDevice(h->device));
  if (!b) {
    DevBuf<double> d;
    QD_ALLOC(h, d, (size_t)N);
    QD_HIP(h, hipMemcpyAsync(d.p, h->FI.p, N * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    int rc = qd_solve_device(h, precision, 1, d.p, N, tol, max_iter, rep);
    QD_HIP(h, hipStreamSynchronize(h->stream));
    return rc;
  }
  if (nrhs < 1 || ldb < N) return set_error(h, QD_ERR_ARG, "nrhs=%d ldb=%d", nrhs, ldb);
  DevBuf<double> d;
  QD_ALLOC(h, d, (size_t)N * nrhs);
  QD_HIP(h, hipMemcpy2D(d.p, N * sizeof(double), b, (size_t)ldb * sizeof(double), N * sizeof(double),
                        nrhs, hipMemcpyHostToDevice));
  int rc = qd_solve_device(h, precision, nrhs, d.p, N, tol, max_iter, rep);
  if (rc < 0) return rc;
  QD_HIP(h, hipStreamSynchronize(h->stream));
  QD_HIP(h, hipMemcpy2D(b, (size_t)ldb * sizeof(double), d.p, N * sizeof(double), N * sizeof(double),
                        nrhs, hipMemcpyDeviceToHost));
  return rc;
}

// io_mod.f90:148-183 (host glue: O(n) work on the solution vector)
int qd_reorganize_solution(qd_handle h, const double* x, double* fi, double* dfi) {
  if (!h || !x || !fi || !dfi) return QD_ERR_ARG;
  if (!h->have_bc || !h->have_mat) return set_error(h, QD_ERR_STATE, "reorganize before qd_set_bc");
  const int n = h->n, ne = h->ne;
  const int* kode = h->hkode.data();
  const double* bc = h->hbc.data();
  const double mu = h->mat.mu;
  if (fi != x) std::memcpy(fi, x, sizeof(double) * 3 * (size_t)n);
  for (int ie = 0; ie < ne; ++ie)
    for (int ld = 0; ld < 3; ++ld)
      for (int k = 0; k < h->hnconec[ie]; ++k) {
        int ik = h->hkconec[ie * 9 + k] - 1;
        dfi[ie * 27 + 3 * k + ld] = (kode[ie * 3 + ld] == 0) ? fi[3 * ik + ld] * mu : bc[ie * 27 + 3 * k + ld];
      }
  for (int in = 0; in < n; ++in)
    for (int ld = 0; ld < 3; ++ld)
      for (int j = 0; j < h->nesim[in]; ++j) {
        int e = h->ke[h->ke_ptr[in] + j] - 1;
        if (kode[e * 3 + ld] == 0) {
          for (int m = 0; m < h->hnconec[e]; ++m)
            if (h->hkconec[e * 9 + m] - 1 == in) {
              fi[3 * in + ld] = bc[e * 27 + 3 * m + ld];
              break;
            }
          break;
        }
      }
  return QD_OK;
}

int qd_internal_points_device(qd_handle h, int l, const double* pts_dev, const double* fi_dev,
                              const double* dfi_dev, int variant, double* u_dev, double* sigma_dev,
                              double* prin_dev, double* vm_dev, long long* counts) {
  if (!h) return QD_ERR_ARG;
  if (l < 0 || (l > 0 && (!pts_dev || !fi_dev || !dfi_dev)) || (variant != 0 && variant != 1))
    return set_error(h, QD_ERR_ARG, "qd_internal_points: l=%d variant=%d", l, variant);
  if (!h->have_mesh || !h->have_mat) return set_error(h, QD_ERR_STATE, "internal points before mesh/material");
  QD_HIP(h, hipSetDevice(h->device));
  if (!h->have_geom) {
    int rc = geometry_run(h);
    if (rc) return rc;
    h->have_geom = true;
  }
  if (l == 0) return QD_OK;
  return somigliana_run(h, l, pts_dev, fi_dev, dfi_dev, variant, u_dev, sigma_dev, prin_dev, vm_dev,
                        counts);
}

int qd_internal_points(qd_handle h, int l, const double* cx, const double* cy, const double* cz,
                       const double* fi, const double* dfi, int variant, double* u_int,
                       double* sigma_int, double* prin, double* vm, long long* counts) {
  if (!h) return QD_ERR_ARG;
  if (l < 0 || (l > 0 && (!cx || !cy || !cz || !fi || !dfi)))
    return set_error(h, QD_ERR_ARG, "qd_internal_points: l=%d or null array", l);
  if (l == 0) return QD_OK;
  QD_HIP(h, hipSetDevice(h->device));
  const size_t L = (size_t)l;
  std::vector<double> pts(3 * L);
  for (size_t i = 0; i < L; ++i) {
    pts[3 * i] = cx[i];
    pts[3 * i + 1] = cy[i];
    pts[3 * i + 2] = cz[i];
  }
  QD_ALLOC(h, h->tmp_pts, 3 * L);
  QD_ALLOC(h, h->tmp_fi, 3 * (size_t)h->n);
  QD_ALLOC(h, h->tmp_dfi, 27 * (size_t)h->ne);
  QD_ALLOC(h, h->tmp_out, 13 * L);
  QD_HIP(h, hipMemcpy(h->tmp_pts.p, pts.data(), 3 * L * sizeof(double), hipMemcpyHostToDevice));
  QD_HIP(h, hipMemcpy(h->tmp_fi.p, fi, 3 * (size_t)h->n * sizeof(double), hipMemcpyHostToDevice));
  QD_HIP(h, hipMemcpy(h->tmp_dfi.p, dfi, 27 * (size_t)h->ne * sizeof(double), hipMemcpyHostToDevice));
  double* o = h->tmp_out.p;
  int rc = qd_internal_points_device(h, l, h->tmp_pts.p, h->tmp_fi.p, h->tmp_dfi.p, variant, o,
                                     o + 3 * L, o + 9 * L, o + 12 * L, counts);
  if (rc) return rc;
  QD_HIP(h, hipStreamSynchronize(h->stream));
  if (u_int) QD_HIP(h, hipMemcpy(u_int, o, 3 * L * sizeof(double), hipMemcpyDeviceToHost));
  if (sigma_int) QD_HIP(h, hipMemcpy(sigma_int, o + 3 * L, 6 * L * sizeof(double), hipMemcpyDeviceToHost));
  if (prin) QD_HIP(h, hipMemcpy(prin, o + 9 * L, 3 * L * sizeof(double), hipMemcpyDeviceToHost));
  if (vm) QD_HIP(h, hipMemcpy(vm, o + 12 * L, L * sizeof(double), hipMemcpyDeviceToHost));
  return QD_OK;
}

int qd_solve_transpose(qd_handle h, int precision, int nrhs, double* c, int ldc, double tol, int max_iter,
                       qd_solve_report* rep) {
  if (!h) return QD_ERR_ARG;
  if (!h->have_sys) return set_error(h, QD_ERR_STATE, "qd_solve_transpose before qd_ghmat8");
  const long long N = 3LL * h->n;
  if (!c || nrhs < 1 || ldc < N) return set_error(h, QD_ERR_ARG, "qd_solve_transpose: nrhs=%d ldc=%d", nrhs, ldc);
  QD_HIP(h, hipSetDevice(h->device));
  DevBuf<double> d;
  QD_ALLOC(h, d, (size_t)N * nrhs);
  QD_HIP(h, hipMemcpy2D(d.p, N * sizeof(double), c, (size_t)ldc * sizeof(double), N * sizeof(double), nrhs,
                        hipMemcpyHostToDevice));
  int rc = qd_solve_transpose_device(h, precision, nrhs, d.p, N, tol, max_iter, rep);
  if (rc < 0) return rc;
  QD_HIP(h, hipStreamSynchronize(h->stream));
  QD_HIP(h, hipMemcpy2D(c, (size_t)ldc * sizeof(double), d.p, N * sizeof(double), N * sizeof(double), nrhs,
                        hipMemcpyDeviceToHost));
  return rc;
}

int qd_ghmat8_vjp(qd_handle h, const double* lambda, const double* fi, const double* dfi, double* gx, double* gy,
                  double* gz, double* gnu) {
  if (!h || !lambda || !fi || !dfi || !gx || !gy || !gz || !gnu) return QD_ERR_ARG;
  if (!h->have_mesh || !h->have_mat || !h->have_bc)
    return set_error(h, QD_ERR_STATE, "qd_ghmat8_vjp needs mesh, material and boundary conditions");
  QD_HIP(h, hipSetDevice(h->device));
  return ghmat8_vjp_run(h, lambda, fi, dfi, gx, gy, gz, gnu);
}

int qd_internal_points_vjp(qd_handle h, int l, const double* cx, const double* cy, const double* cz,
                           const double* fi, const double* dfi, int variant, const double* g_u,
                           const double* g_sigma, double* g_fi, double* g_dfi, double* gx, double* gy, double* gz,
                           double* gnu) {
  if (!h || !fi || !dfi || !g_fi || !g_dfi || !gx || !gy || !gz || !gnu) return QD_ERR_ARG;
  if (l < 0 || (l > 0 && (!cx || !cy || !cz)) || (variant != 0 && variant != 1))
    return set_error(h, QD_ERR_ARG, "qd_internal_points_vjp: l=%d variant=%d", l, variant);
  if (!h->have_mesh || !h->have_mat)
    return set_error(h, QD_ERR_STATE, "qd_internal_points_vjp before mesh/material");
  QD_HIP(h, hipSetDevice(h->device));
  const size_t n = (size_t)h->n, ne = (size_t)h->ne;
  if (l == 0) {
    std::memset(g_fi, 0, 3 * n * sizeof(double));
    std::memset(g_dfi, 0, 27 * ne * sizeof(double));
    std::memset(gx, 0, n * sizeof(double));
    std::memset(gy, 0, n * sizeof(double));
    std::memset(gz, 0, n * sizeof(double));
    *gnu = 0.0;
    return QD_OK;
  }
  std::vector<double> pts(3 * (size_t)l);
  for (int i = 0; i < l; ++i) {
    pts[3 * (size_t)i] = cx[i];
    pts[3 * (size_t)i + 1] = cy[i];
    pts[3 * (size_t)i + 2] = cz[i];
  }
  return somigliana_vjp_run(h, l, pts.data(), fi, dfi, variant, g_u, g_sigma, g_fi, g_dfi, gx, gy, gz, gnu);
}

int qd_fp16_update(qd_handle h, long long M, long long N, int K, const void* A16, const void* B16,
                   const float* colscale, float* C, long long ldc) {
  if (!h) return QD_ERR_ARG;
  if (M < 0 || N < 0 || K <= 0 || K % 32 != 0 || ldc < N || (M > 0 && N > 0 && (!A16 || !B16 || !colscale || !C)))
    return set_error(h, QD_ERR_ARG, "qd_fp16_update: M=%lld N=%lld K=%d ldc=%lld", M, N, K, ldc);
  QD_HIP(h, hipSetDevice(h->device));
  return fp16_update(h, M, N, K, A16, B16, colscale, C, ldc);
}

int qd_memory_usage(long long* current_bytes, long long* peak_bytes, int reset_peak) {
  MemStats& m = mem_stats();
  if (current_bytes) *current_bytes = m.current.load();
  if (peak_bytes) *peak_bytes = m.peak.load();
  if (reset_peak) m.peak.store(m.current.load());
  return QD_OK;
}

int qd_profile_enable(qd_handle h, unsigned mask) {
  if (!h) return QD_ERR_ARG;
  h->prof_mask = mask;
  return QD_OK;
}

int qd_profile_read(qd_handle h, int kind, long long* launches, double* ms, double* work) {
  if (!h || kind < 0 || kind >= QD_PROF_KINDS) return QD_ERR_ARG;
  QD_HIP(h, hipSetDevice(h->device));
  QD_HIP(h, hipStreamSynchronize(h->stream));
  for (auto& r : h->prof_pending) {
    float t = 0.0f;
    if (hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
      h->prof_launches[r.kind] += 1;
      h->prof_ms[r.kind] += t;
      h->prof_work[r.kind] += r.work;
    }
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  h->prof_pending.clear();
  if (launches) *launches = h->prof_launches[kind];
  if (ms) *ms = h->prof_ms[kind];
  if (work) *work = h->prof_work[kind];
  // reading resets the class
  h->prof_launches[kind] = 0;
  h->prof_ms[kind] = 0.0;
  h->prof_work[kind] = 0.0;
  return QD_OK;
}

int qd_boundary_stresses(qd_handle h, const double* fi, const double* dfi, double* sigma_elem,
                         double* sigma_node, double* prin, double* vm) {
  if (!h || !fi || !dfi) return QD_ERR_ARG;
  if (!h->have_mesh || !h->have_mat) return set_error(h, QD_ERR_STATE, "boundary stresses before mesh/material");
  QD_HIP(h, hipSetDevice(h->device));
  const size_t n = h->n, ne = h->ne;
  DevBuf<double> dfi_, ddfi, out;
  QD_ALLOC(h, dfi_, 3 * n);
  QD_ALLOC(h, ddfi, 27 * ne);
  QD_ALLOC(h, out, 54 * ne + 10 * n);
  QD_HIP(h, hipMemcpy(dfi_.p, fi, 3 * n * sizeof(double), hipMemcpyHostToDevice));
  QD_HIP(h, hipMemcpy(ddfi.p, dfi, 27 * ne * sizeof(double), hipMemcpyHostToDevice));
  double* se = out.p;
  double* sn = se + 54 * ne;
  double* pr = sn + 6 * n;
  double* v = pr + 3 * n;
  int rc = boundary_stress_run(h, dfi_.p, ddfi.p, se, sn, pr, v);
  if (rc) return rc;
  QD_HIP(h, hipStreamSynchronize(h->stream));
  if (sigma_elem) QD_HIP(h, hipMemcpy(sigma_elem, se, 54 * ne * sizeof(double), hipMemcpyDeviceToHost));
  if (sigma_node) QD_HIP(h, hipMemcpy(sigma_node, sn, 6 * n * sizeof(double), hipMemcpyDeviceToHost));
  if (prin) QD_HIP(h, hipMemcpy(prin, pr, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
  if (vm) QD_HIP(h, hipMemcpy(vm, v, n * sizeof(double), hipMemcpyDeviceToHost));
  return QD_OK;
}

}  // extern "C"
