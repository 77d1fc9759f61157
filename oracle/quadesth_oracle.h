/*
 * quadesth_oracle.h -- CPU FP64 restatement of QUADESTH's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP library is
 * compared against; only the tests, the smoke check and the CPU-baseline leg
 * of the benchmark may load it.  The product path never calls into it.
 *
 * Array conventions are the reference's module globals (data_mod.f90:38-52),
 * Fortran column-major, with 0-based C offsets:
 *   x,y,z[n]; nconec[ne]; kconec[e*9+m] (1-based node numbers);
 *   kode[e*3+ld]; bc[e*27 + 3*k + ld]; nesim[n]; ke[node*cap + i] (1-based
 *   element numbers, cap = qo_ke_cap(), 8 by default); A[col*N + row], N = 3n; FI[N]; DFI[e*27 + 3*k + ld];
 *   Gauss tables g[e*36 + k] for eta1..3, xja, xg1..3 and ff[(e*36+k)*9 + m].
 */
#ifndef QUADESTH_ORACLE_H
#define QUADESTH_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double mu, nu, young;
  double c3, c4, c5, c6;
} qo_material_t;

void qo_material(double mu, double nu, qo_material_t* m);

void qo_shape(int ncone, double g1, double g2, double ff[9], double fd1[9], double fd2[9]);

void qo_etaxja(const double xj1[9], const double xj2[9], const double xj3[9],
               double eta1[36], double eta2[36], double eta3[36], double xja[36],
               double xg1[36], double xg2[36], double xg3[36], double ff[9 * 36], int ncone);

void qo_precompute(int ne, const int* nconec, const int* kconec,
                   const double* x, const double* y, const double* z,
                   double* eta1, double* eta2, double* eta3, double* xja,
                   double* xg1, double* xg2, double* xg3, double* ff);

void qo_funkelvin(const qo_material_t* m, double xp, double yp, double zp, const double eta[3],
                  double xg, double yg, double zg, double u[9], double t[9]);

void qo_extin8(const qo_material_t* m, const double* eta1, const double* eta2, const double* eta3,
               const double* xja, const double* xg1, const double* xg2, const double* xg3,
               const double* ff, const double xi[3], double h[81], double g[81], int ncone);

void qo_locin8(const qo_material_t* m, const double xj1[9], const double xj2[9], const double xj3[9],
               double h[81], double g[81], int nodo, int ncone);

/* Width of ke per node (default 8, the reference's ke(8, n)); ke[node*cap + i]. */
void qo_set_ke_cap(int cap);
int qo_ke_cap(void);
void qo_adjacency(int n, int ne, const int* nconec, const int* kconec, int* nesim, int* ke);

void qo_ghmat8(int n, int ne, const double* x, const double* y, const double* z,
               const int* nconec, const int* kconec, const int* kode, const double* bc,
               const int* nesim, const int* ke, const qo_material_t* m,
               const double* eta1, const double* eta2, const double* eta3, const double* xja,
               const double* xg1, const double* xg2, const double* xg3, const double* ff,
               double* A, double* FI);

/* ghmat8 restricted to the collocation nodes listed (0-based): rows
 * 3*nodes[i]+jj of A into Arows[(3i+jj)*3n + col], FI into FIrows[3i+jj]. */
void qo_ghmat8_rows(int n, int ne, const double* x, const double* y, const double* z,
                    const int* nconec, const int* kconec, const int* kode, const double* bc,
                    const int* nesim, const int* ke, const qo_material_t* m,
                    const double* eta1, const double* eta2, const double* eta3, const double* xja,
                    const double* xg1, const double* xg2, const double* xg3, const double* ff,
                    int nsel, const int* nodes, double* Arows, double* FIrows);

/* Element integrals of one (collocation node, element) pair as ghmat8 sees
 * them (regular or singular); h, g are 3x27 column-major. */
void qo_element_pair(int in, int iej, const double* x, const double* y, const double* z,
                     const int* nconec, const int* kconec, const qo_material_t* m,
                     const double* eta1, const double* eta2, const double* eta3, const double* xja,
                     const double* xg1, const double* xg2, const double* xg3, const double* ff,
                     double h[81], double g[81]);

int qo_dgesv(int n, int nrhs, double* a, int lda, int* ipiv, double* b, int ldb);

void qo_reorganize(int n, int ne, const int* nconec, const int* kconec, const int* kode,
                   const double* bc, const int* nesim, const int* ke, double mu,
                   double* FI, double* DFI);

void qo_gauss_legendre(int npts, double* xg, double* wg);

void qo_principal_vm(const double sig[6], double prin[3], double* vm);

/* variant 0 = as the reference (stress_mod.f90:838-843); 1 = textbook sign of
 * the antisymmetric T* term in the displacement identity. */
void qo_internal_points(int n, int ne, const double* x, const double* y, const double* z,
                        const int* nconec, const int* kconec, const qo_material_t* m,
                        const double* eta1, const double* eta2, const double* eta3, const double* xja,
                        const double* xg1, const double* xg2, const double* xg3, const double* ff,
                        const double* FI, const double* DFI, int l,
                        const double* cx, const double* cy, const double* cz, int variant,
                        double* u_int, double* sigma_int, double* prin, double* vm,
                        long long* counts /* [2]: far, near GP evaluations, or NULL */);

void qo_boundary_stresses(int n, int ne, const double* x, const double* y, const double* z,
                          const int* nconec, const int* kconec, const int* nesim, const int* ke,
                          const qo_material_t* m, const double* FI, const double* DFI,
                          double* sigma_elem /* 6*9*ne */, double* sigma_node /* 6*n */,
                          double* prin /* 3*n */, double* vm /* n */);

#ifdef __cplusplus
}
#endif
#endif
