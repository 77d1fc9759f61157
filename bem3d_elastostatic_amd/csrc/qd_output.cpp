// qd_output.cpp -- the reference's text outputs, written from host arrays:
// the deck echo that input_data writes (io_mod.f90:44-140), output_data
// (io_mod.f90:191-228), output_boundary_stresses (stress_mod.f90:412-468),
// output_internal_points (stress_mod.f90:892-936) and save_disp_trac
// (io_mod.f90:230-255).  Each record follows the reference's FORMAT
// statement, so the files read back with the same column positions (and
// posproc.py-style scripts parse them unchanged).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "qd_internal.h"

namespace {

std::string blanks(int n) { return std::string(n > 0 ? n : 0, ' '); }

std::string right(const std::string& s, int w) {
  if ((int)s.size() > w) return std::string(w, '*');  // the field overflows
  return blanks(w - (int)s.size()) + s;
}

// Iw
std::string fI(long long v, int w) { return right(std::to_string(v), w); }

std::string nonfinite(double v, int w) {
  if (std::isnan(v)) return right("NaN", w);
  const char* s = v < 0 ? (w >= 9 ? "-Infinity" : "-Inf") : (w >= 8 ? "Infinity" : "Inf");
  return right(s, w);
}

// Ew.d with the default scale factor: [-]0.d1..dd E+xx, the d digits rounded
// to nearest from the binary value.  Exponents beyond two digits drop the E
// (0.1234567+100), as the standard's Ew.d form does.
std::string fE(double v, int w, int d) {
  if (!std::isfinite(v)) return nonfinite(v, w);
  const bool neg = std::signbit(v);
  std::string mant, sexp;
  int e10 = 0;
  if (v == 0.0) {
    mant = std::string(d, '0');
  } else {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.*E", d - 1, std::fabs(v));  // D.DDDDDDE+XX
    const char* ep = std::strchr(buf, 'E');
    mant.assign(1, buf[0]);
    if (d > 1) mant.append(buf + 2, (size_t)(ep - (buf + 2)));
    e10 = std::atoi(ep + 1) + 1;
  }
  char eb[16];
  if (std::abs(e10) <= 99)
    std::snprintf(eb, sizeof(eb), "E%c%02d", e10 < 0 ? '-' : '+', std::abs(e10));
  else
    std::snprintf(eb, sizeof(eb), "%c%03d", e10 < 0 ? '-' : '+', std::abs(e10));
  sexp = eb;
  std::string body = "0." + mant + sexp;
  if ((int)body.size() + (neg ? 1 : 0) > w) body = "." + mant + sexp;  // the optional zero goes first
  return right((neg ? "-" : "") + body, w);
}

// Fw.d
std::string fF(double v, int w, int d) {
  if (!std::isfinite(v)) return nonfinite(v, w);
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.*f", d, v);
  std::string s = buf;
  if (d == 0) s += ".";
  if (std::signbit(v) && s[0] != '-') s = "-" + s;
  if ((int)s.size() > w && (s.rfind("0.", 0) == 0 || s.rfind("-0.", 0) == 0)) s.erase(s[0] == '-' ? 1 : 0, 1);
  return right(s, w);
}

// Gw.d for a real: F editing with four trailing blanks when the value,
// rounded to d significant digits, lies in [0.1, 10^d) (or is zero), else
// Ew.d.
std::string fG(double v, int w, int d) {
  if (!std::isfinite(v)) return nonfinite(v, w);
  if (v == 0.0) return fF(v, w - 4, d - 1) + blanks(4);
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.*E", d - 1, std::fabs(v));
  const int k = std::atoi(std::strchr(buf, 'E') + 1) + 1;  // digits before the point
  if (k >= 0 && k <= d) return fF(v, w - 4, d - k) + blanks(4);
  return fE(v, w, d);
}

const std::string kStars = " " + std::string(79, '*');

struct Out {
  FILE* f = nullptr;
  explicit Out(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
  ~Out() {
    if (f) std::fclose(f);
  }
  void rec(const std::string& s) { std::fputs((s + "\n").c_str(), f); }
  void empty(int k = 1) {
    for (int i = 0; i < k; ++i) std::fputs("\n", f);
  }
  bool ok() const { return f && !std::ferror(f); }
};

int io_error(qd_context* h, const char* what, const char* path) {
  return qd::set_error(h, QD_ERR_IO, "%s: cannot write %s", what, path ? path : "(null)");
}

// ke(1, node): the first element, in ascending order, that holds the node.
int first_element(const qd_context* h, int node0) { return h->ke[h->ke_ptr[node0]] - 1; }

}  // namespace

using namespace qd;

extern "C" {

int qd_write_input_echo(qd_handle h, const char* path, const char* title, int l, const double* cx,
                        const double* cy, const double* cz, int ntrack) {
  if (!h || !path || (l > 0 && (!cx || !cy || !cz))) return QD_ERR_ARG;
  if (!h->have_mesh || !h->have_mat || !h->have_bc)
    return set_error(h, QD_ERR_STATE, "qd_write_input_echo needs mesh, material and boundary conditions");
  Out o(path, "w");
  if (!o.f) return io_error(h, "qd_write_input_echo", path);
  const int n = h->n, ne = h->ne;
  // io_mod.f90:44-46: the title as read with '(A)' into character(len=80)
  std::string t = title ? title : "";
  if (t.size() > 80) t.resize(80);
  while (!t.empty() && t.back() == ' ') t.pop_back();
  o.rec(kStars);
  o.rec(t);
  // FORMAT 300 (io_mod.f90:66-71)
  o.empty(2);
  o.rec("  DATA");
  o.empty();
  o.rec("  NUMBER OF BOUNDARY NODES    =" + fI(n, 5));
  o.rec("  NUMBER OF BOUNDARY ELEMENTS =" + fI(ne, 5));
  o.rec("  NUMBER OF INTERNAL POINTS   =" + fI(l, 5));
  o.rec("  SHEAR MODULUS (MU)          =" + fG(h->mat.mu, 14, 7));
  o.rec("  YOUNG MODULUS (E)           =" + fG(h->mat.young, 14, 7));
  o.rec("  POISSON RATIO (NU)          =" + fG(h->mat.nu, 14, 7));
  // FORMAT 500 and 240 (io_mod.f90:74-84)
  o.empty(2);
  o.rec("  BOUNDARY NODES COORDINATES");
  o.empty(2);
  o.rec("    NODE" + blanks(12) + "X" + blanks(18) + "Y" + blanks(18) + "Z");
  o.empty();
  for (int i = 0; i < n; ++i)
    o.rec(blanks(5) + fI(i + 1, 5) + blanks(5) + fE(h->hx[i], 14, 7) + blanks(5) + fE(h->hy[i], 14, 7) +
          blanks(5) + fE(h->hz[i], 14, 7));
  // FORMAT 600 and 340 (io_mod.f90:87-93)
  o.empty(2);
  o.rec("  ELEMENT CONNECTIVITIES");
  o.empty(2);
  o.rec("  ELEMENT NO.CON." + blanks(24) + "NODES");
  o.empty();
  for (int e = 0; e < ne; ++e) {
    std::string s = blanks(4) + fI(e + 1, 5) + blanks(4) + fI(h->hnconec[e], 2) + blanks(2);
    for (int j = 0; j < h->hnconec[e]; ++j) s += blanks(3) + fI(h->hkconec[e * 9 + j], 5);
    o.rec(s);
  }
  // FORMAT 800 and 950-970 (io_mod.f90:96-113)
  o.empty(2);
  o.rec("  BOUNDARY CONDITIONS");
  o.empty();
  o.rec(" ELEMEN CODE" + blanks(22) + "PRESCRIBED VALUES");
  o.empty();
  const char dirs[3] = {'X', 'Y', 'Z'};
  for (int e = 0; e < ne; ++e)
    for (int k = 0; k < 3; ++k) {
      std::string s = (k == 0) ? " " + fI(e + 1, 5) + blanks(4) : blanks(8);
      s += fI(h->hkode[e * 3 + k], 2) + " " + dirs[k];
      for (int j = 0; j < h->hnconec[e]; ++j) s += " " + fG(h->hbc[e * 27 + 3 * j + k], 12, 5);
      o.rec(s);
    }
  // FORMAT 1000 and 240 (io_mod.f90:116-124)
  if (l > 0) {
    o.empty(2);
    o.rec("  COORDINATES OF INTERNAL POINTS");
    o.empty();
    o.rec("    POINT" + blanks(11) + "X" + blanks(18) + "Y" + blanks(17) + " Z");
    o.empty();
    for (int i = 0; i < l; ++i)
      o.rec(blanks(5) + fI(i + 1, 5) + blanks(5) + fE(cx[i], 14, 7) + blanks(5) + fE(cy[i], 14, 7) + blanks(5) +
            fE(cz[i], 14, 7));
  }
  // io_mod.f90:140
  o.rec(" TRACKED NODE = " + fI(ntrack, 6));
  return o.ok() ? QD_OK : io_error(h, "qd_write_input_echo", path);
}

int qd_output_data(qd_handle h, const char* path, const double* fi, const double* dfi, int l, const double* dsol) {
  if (!h || !path || !fi || !dfi || l < 0 || (l > 0 && !dsol)) return QD_ERR_ARG;
  if (!h->have_mesh) return set_error(h, QD_ERR_STATE, "qd_output_data before qd_set_mesh");
  Out o(path, "a");
  if (!o.f) return io_error(h, "qd_output_data", path);
  const int n = h->n, ne = h->ne;
  o.rec(kStars);
  o.rec(" QUADESTH: 3D ELASTOSTATIC ANALYSIS RESULTS");
  o.rec(kStars);
  o.rec(" BOUNDARY NODES DISPLACEMENTS");
  o.rec(" NODE           DISPLACEMENT X          DISPLACEMENT Y          DISPLACEMENT Z");
  auto f200 = [&](long long i, const double* v) {  // FORMAT 200 (1X, I5, 3(5X, E18.8))
    o.rec(" " + fI(i, 5) + blanks(5) + fE(v[0], 18, 8) + blanks(5) + fE(v[1], 18, 8) + blanks(5) + fE(v[2], 18, 8));
  };
  for (int i = 0; i < n; ++i) f200(i + 1, fi + 3 * (size_t)i);
  o.empty(2);  // FORMAT 110
  o.rec("  BOUNDARY ELEMENTS TRACTIONS");
  o.empty();
  o.rec(" ELEMENT  NODE" + blanks(8) + "TRACTION X" + blanks(14) + "TRACTION Y" + blanks(14) + "TRACTION Z");
  o.empty();
  for (int e = 0; e < ne; ++e) {
    o.rec(" " + fI(e + 1, 5));
    for (int k = 0; k < h->hnconec[e]; ++k) {  // FORMAT 130 (10X, I5, 3(5X, E18.8))
      const double* t = dfi + (size_t)e * 27 + 3 * k;
      o.rec(blanks(10) + fI(h->hkconec[e * 9 + k], 5) + blanks(5) + fE(t[0], 18, 8) + blanks(5) + fE(t[1], 18, 8) +
            blanks(5) + fE(t[2], 18, 8));
    }
  }
  if (l > 0) {
    o.empty(2);  // FORMAT 310
    o.rec("  INTERNAL POINTS DISPLACEMENTS");
    o.empty();
    o.rec("POINT" + blanks(7) + "DISPLACEMENT X" + blanks(10) + "DISPLACEMENT Y" + blanks(10) + "DISPLACEMENT Z");
    o.empty();
    for (int k = 0; k < l; ++k) f200(k + 1, dsol + 3 * (size_t)k);
  }
  o.rec(kStars);
  return o.ok() ? QD_OK : io_error(h, "qd_output_data", path);
}

int qd_output_boundary_stresses(qd_handle h, const char* path, const char* stresses_path, const double* sigma_elem,
                                const double* sigma_node, const double* prin, const double* vm) {
  if (!h || !sigma_elem || !sigma_node || !prin || !vm || (!path && !stresses_path)) return QD_ERR_ARG;
  if (!h->have_mesh) return set_error(h, QD_ERR_STATE, "qd_output_boundary_stresses before qd_set_mesh");
  const int n = h->n, ne = h->ne;
  auto ten = [&](int i) {  // 10(1X, E15.7): sigma(6), von Mises, principal(3)
    std::string s;
    for (int k = 0; k < 6; ++k) s += " " + fE(sigma_node[6 * (size_t)i + k], 15, 7);
    s += " " + fE(vm[i], 15, 7);
    for (int k = 0; k < 3; ++k) s += " " + fE(prin[3 * (size_t)i + k], 15, 7);
    return s;
  };
  if (path) {
    Out o(path, "a");
    if (!o.f) return io_error(h, "qd_output_boundary_stresses", path);
    o.rec(kStars);
    o.rec(" BOUNDARY NODES STRESSES (NODAL AVERAGED)");
    o.rec(kStars);
    o.rec(" NODE      SIGMA_XX        SIGMA_YY        SIGMA_ZZ        SIGMA_XY        SIGMA_YZ        SIGMA_ZX"
          "       VON_MISES         SIGMA_1         SIGMA_2         SIGMA_3");
    o.empty();
    for (int i = 0; i < n; ++i) o.rec(" " + fI(i + 1, 5) + ten(i));  // FORMAT 110
    o.empty(2);
    o.rec(kStars);
    o.rec(" BOUNDARY ELEMENTS NODAL STRESSES");
    o.rec(kStars);
    o.rec(" ELEMENT  NODE      SIGMA_XX        SIGMA_YY        SIGMA_ZZ        SIGMA_XY        SIGMA_YZ        SIGMA_ZX");
    o.empty();
    for (int e = 0; e < ne; ++e) {
      o.rec(" " + fI(e + 1, 5));
      for (int k = 0; k < h->hnconec[e]; ++k) {  // FORMAT 210 (10X, I5, 6(1X, E15.7))
        std::string s = blanks(10) + fI(h->hkconec[e * 9 + k], 5);
        for (int c = 0; c < 6; ++c) s += " " + fE(sigma_elem[(size_t)e * 54 + 6 * k + c], 15, 7);
        o.rec(s);
      }
    }
    if (!o.ok()) return io_error(h, "qd_output_boundary_stresses", path);
  }
  if (stresses_path) {
    Out o(stresses_path, "w");
    if (!o.f) return io_error(h, "qd_output_boundary_stresses", stresses_path);
    o.rec("# BOUNDARY NODES STRESSES (QUADESTH 3D ELASTOSTATICS)");
    o.rec("# Col 1: Node ID");
    o.rec("# Col 2-4: X, Y, Z coordinates");
    o.rec("# Col 5-10: Sigma_xx, Sigma_yy, Sigma_zz, Sigma_xy, Sigma_yz, Sigma_zx");
    o.rec("# Col 11: Von Mises Equivalent Stress");
    o.rec("# Col 12-14: Principal Stresses (Sigma_1 >= Sigma_2 >= Sigma_3)");
    o.rec("# " + std::string(140, '-'));
    for (int i = 0; i < n; ++i)  // FORMAT 300 (I6, 3(1X, E15.7), 10(1X, E15.7))
      o.rec(fI(i + 1, 6) + " " + fE(h->hx[i], 15, 7) + " " + fE(h->hy[i], 15, 7) + " " + fE(h->hz[i], 15, 7) + ten(i));
    if (!o.ok()) return io_error(h, "qd_output_boundary_stresses", stresses_path);
  }
  return QD_OK;
}

int qd_output_internal_points(qd_handle h, const char* path, const char* points_path, int l, const double* cx,
                              const double* cy, const double* cz, const double* u_int, const double* sigma_int,
                              const double* prin, const double* vm) {
  if (!h || l < 0 || (!path && !points_path)) return QD_ERR_ARG;
  if (l == 0) return QD_OK;  // stress_mod.f90:897
  if (!cx || !cy || !cz || !u_int || !sigma_int || !prin || !vm) return QD_ERR_ARG;
  auto ten = [&](int i) {  // sigma(6), von Mises, principal(3)
    std::string s;
    for (int k = 0; k < 6; ++k) s += " " + fE(sigma_int[6 * (size_t)i + k], 15, 7);
    s += " " + fE(vm[i], 15, 7);
    for (int k = 0; k < 3; ++k) s += " " + fE(prin[3 * (size_t)i + k], 15, 7);
    return s;
  };
  if (path) {
    Out o(path, "a");
    if (!o.f) return io_error(h, "qd_output_internal_points", path);
    o.empty(2);
    o.rec(kStars);
    o.rec(" INTERNAL POINTS STRESSES");
    o.rec(kStars);
    o.rec(" POINT     SIGMA_XX        SIGMA_YY        SIGMA_ZZ        SIGMA_XY        SIGMA_YZ        SIGMA_ZX"
          "       VON_MISES         SIGMA_1         SIGMA_2         SIGMA_3");
    o.empty();
    for (int i = 0; i < l; ++i) o.rec(" " + fI(i + 1, 5) + ten(i));  // FORMAT 110
    if (!o.ok()) return io_error(h, "qd_output_internal_points", path);
  }
  if (points_path) {
    Out o(points_path, "w");
    if (!o.f) return io_error(h, "qd_output_internal_points", points_path);
    o.rec("# INTERNAL POINTS DISPLACEMENTS AND STRESSES (QUADESTH 3D ELASTOSTATICS)");
    o.rec("# Total Internal Points: " + fI(l, 6));
    o.rec("# Col 1: Point ID");
    o.rec("# Col 2-4: X, Y, Z coordinates");
    o.rec("# Col 5-7: u_x, u_y, u_z displacements");
    o.rec("# Col 8-13: Sigma_xx, Sigma_yy, Sigma_zz, Sigma_xy, Sigma_yz, Sigma_zx");
    o.rec("# Col 14: Von Mises Equivalent Stress");
    o.rec("# Col 15-17: Principal Stresses (Sigma_1 >= Sigma_2 >= Sigma_3)");
    o.rec("# " + std::string(140, '-'));
    for (int i = 0; i < l; ++i) {  // FORMAT 300 (I6, 3(1X, E15.7), 3(1X, E15.7), 6(1X, E15.7), 4(1X, E15.7))
      std::string s = fI(i + 1, 6) + " " + fE(cx[i], 15, 7) + " " + fE(cy[i], 15, 7) + " " + fE(cz[i], 15, 7);
      for (int k = 0; k < 3; ++k) s += " " + fE(u_int[3 * (size_t)i + k], 15, 7);
      o.rec(s + ten(i));
    }
    if (!o.ok()) return io_error(h, "qd_output_internal_points", points_path);
  }
  return QD_OK;
}

int qd_save_disp_trac(qd_handle h, const char* path, int append, int ntrack, const double* fi, const double* dfi) {
  if (!h || !path || !fi || !dfi) return QD_ERR_ARG;
  if (!h->have_mesh) return set_error(h, QD_ERR_STATE, "qd_save_disp_trac before qd_set_mesh");
  if (ntrack < 1 || ntrack > h->n) return set_error(h, QD_ERR_ARG, "tracked node %d out of 1..%d", ntrack, h->n);
  // io_mod.f90:235-250: the node's slot in its first element
  const int ie = first_element(h, ntrack - 1);
  int kloc = 0;
  for (int k = 0; k < h->hnconec[ie]; ++k)
    if (h->hkconec[ie * 9 + k] == ntrack) {
      kloc = k;
      break;
    }
  Out o(path, append ? "a" : "w");
  if (!o.f) return io_error(h, "qd_save_disp_trac", path);
  const double* u = fi + 3 * (size_t)(ntrack - 1);
  const double* t = dfi + (size_t)ie * 27 + 3 * kloc;
  std::string s;  // FORMAT 600 (6(2X, E18.8))
  for (int k = 0; k < 3; ++k) s += blanks(2) + fE(u[k], 18, 8);
  for (int k = 0; k < 3; ++k) s += blanks(2) + fE(t[k], 18, 8);
  o.rec(s);
  return o.ok() ? QD_OK : io_error(h, "qd_save_disp_trac", path);
}

}  // extern "C"
