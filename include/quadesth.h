/*
 * quadesth.h -- C ABI of libquadesth, the MI355X (gfx950) implementation of
 * QUADESTH's hot path: the ghmat8 collocation assembly, the dense DGESV
 * solve and the Somigliana internal-point integrals.
 *
 * Every entry point replaces one routine of the reference (cited file:line
 * into the QUADESTH sources).  Arrays use the reference's own layouts (the
 * data_mod.f90:38-52 module globals, Fortran column-major, element-node
 * arrays (27, ne) with dof 3*(k-1)+ld, kconec(9, ne) with 1-based node
 * numbers), so a Fortran caller passes its module arrays unchanged through
 * ISO_C_BINDING and a Python caller passes numpy/torch buffers.
 *
 * Pointers named *_dev are device (HBM) pointers; all others are host
 * pointers.  Sizes are int (the reference's default integer); matrix
 * leading dimensions and offsets are 64-bit internally (76,806^2 > 2^31).
 *
 * Return value: 0 = success; > 0 = LAPACK-style singular pivot (1-based);
 * < 0 = error (see QD_ERR_*; qd_last_error gives the message).  The
 * reference has no error returns: its DGESV failure prints and stops
 * (solver_mod.f90:30-33); a Fortran shim keeps that behaviour on non-zero.
 */
#ifndef QUADESTH_H
#define QUADESTH_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qd_context* qd_handle;

enum {
  QD_OK = 0,
  QD_ERR_ARG = -1,     /* bad argument */
  QD_ERR_HIP = -2,     /* HIP runtime error */
  QD_ERR_STATE = -3,   /* called before the phase it depends on */
  QD_ERR_MESH = -4,    /* invalid mesh: nconec not 9 or 6, a node number out
                          of range, a node in no element */
  QD_ERR_NOMEM = -5,   /* device allocation failed */
  QD_ERR_NOCONV = -6,  /* iterative refinement missed its tolerance */
  QD_ERR_IO = -7       /* an output file could not be written */
};

/* Precision of the factorisation used by qd_solve.  Residuals of the
 * iterative refinement are always formed with the FP64 matrix. */
enum {
  QD_FP64 = 0,      /* FP64 blocked LU (MFMA f64 trailing updates) */
  QD_FP32_IR = 1,   /* FP32 LU (MFMA f32), FP64 iterative refinement */
  QD_FP16_IR = 2    /* FP32 panel + FP16 MFMA trailing updates with
                       per-block scaling, FP64 refinement (GMRES fallback) */
};

/* Arithmetic of the regular element integrals of ghmat8 (extin8,
 * assembly_mod.f90:339-384).  A pair is far when the collocation node lies
 * more than two element diameters (twice the largest node-to-centroid
 * distance) from the element's centroid; far pairs may run in FP32 (the
 * Kelvin kernels and the Gauss sums; r itself in FP64).  Near and adjacent
 * pairs, the singular elements (locin8), the rigid-body diagonal and every
 * sum into A and FI stay FP64. */
enum {
  QD_ASSEMBLY_FP64 = 0,            /* every integral FP64 (the reference's) */
  QD_ASSEMBLY_FAR_FP32 = 1,        /* A itself with FP32 far pairs: entries
                                      off by ~1e-7 relative, so every solve
                                      solves that A */
  QD_ASSEMBLY_FAR_FP32_FACTOR = 2  /* A in FP64; the FP32/FP16 factorisations
                                      factor a second matrix assembled with
                                      FP32 far pairs, while the refinement's
                                      residual keeps the FP64 A (solutions as
                                      with QD_ASSEMBLY_FP64) */
};

/* Somigliana displacement identity variant (stress_mod.f90:838-843). */
enum {
  QD_SOMIGLIANA_REFERENCE = 0,  /* the reference's sign of the T* term */
  QD_SOMIGLIANA_CORRECTED = 1   /* the sign the assembly kernel uses
                                   (assembly_mod.f90:646-647) */
};

typedef struct {
  int precision;        /* QD_FP64 / QD_FP32_IR / QD_FP16_IR */
  int info;             /* LAPACK-style info of the factorisation */
  int iterations;       /* refinement sweeps taken (0 for QD_FP64) */
  int converged;        /* 1 when the residual met tol (or QD_FP64) */
  int gmres;            /* 1 when the GMRES fallback was used */
  int pivoting;         /* mixed precisions: 1 when the factors kept their
                           diagonal pivots (every multiplier |l| <= 1, so
                           partial pivoting's choice), 0 when they use
                           tournament pivoting */
  double residual;      /* max over rhs of ||b-Ax||_inf/(||A||_inf||x||_inf+||b||_inf) */
  double t_factor_ms;   /* device time of the factorisation */
  double t_solve_ms;    /* device time of substitutions + refinement */
} qd_solve_report;

/* ---- lifecycle ----------------------------------------------------------- */
int qd_create(qd_handle* out, int device);
void qd_destroy(qd_handle h);
/* Message of the last failed call on h; with h == NULL, of the last failed
 * qd_create on the calling thread. */
const char* qd_last_error(qd_handle h);
/* Work is queued on this HIP stream (default: the null stream). */
int qd_set_stream(qd_handle h, void* hip_stream);
int qd_synchronize(qd_handle h);

/* ---- problem data ---------------------------------------------------------
 * Replaces input_data's fill of the data_mod globals (io_mod.f90:36-146) and
 * the nesim/ke adjacency (io_mod.f90:124-135).  Copies to the device. */
int qd_set_mesh(qd_handle h, int n, int ne, const double* x, const double* y, const double* z,
                const int* nconec, const int* kconec);
/* update_material_constants (data_mod.f90:97-108): mu is the shear modulus
 * the deck carries (rge), nu Poisson's ratio. */
int qd_set_material(qd_handle h, double mu, double nu);
/* kode(3*ne): 0 = displacement prescribed, 1 = traction prescribed;
 * bc(27, ne).  The system matrix depends on kode only; bc enters FI. */
int qd_set_bc(qd_handle h, const int* kode, const double* bc);

/* ---- hot path ------------------------------------------------------------ */
/* precompute_element_geometry (assembly_mod.f90:28-51).  Optionally returns
 * the all_* tables (each (36, ne), ff (9, 36, ne)); pass NULL to skip. */
int qd_precompute_element_geometry(qd_handle h, double* eta1, double* eta2, double* eta3,
                                   double* xja, double* xg1, double* xg2, double* xg3, double* ff);

/* ghmat8 (assembly_mod.f90:53-223): assembles A (3n x 3n) and FI (3n) into
 * device memory owned by the handle. */
int qd_ghmat8(qd_handle h);
/* QD_ASSEMBLY_* for the next qd_ghmat8 (no reference counterpart: the
 * reference integrates in FP64 only). */
int qd_set_assembly_precision(qd_handle h, int mode);
/* FI only, for another set of prescribed values with the same kode
 * (assembly_mod.f90:126-218 restricted to FI); A is untouched. */
int qd_ghmat8_rhs(qd_handle h, const double* bc, double* fi /* host, 3n */);
/* Copies of the assembled system (host, column-major with leading dim lda). */
int qd_get_system(qd_handle h, double* A /* nullable */, int lda, double* FI /* nullable */);
/* Device views of the handle's A (column-major, leading dimension *lda) and FI. */
int qd_system_device(qd_handle h, double** A_dev, long long* lda, double** FI_dev);

/* solve_linear_system (solver_mod.f90:17-36) on the device-resident A.
 * b: host (N x nrhs, column-major, leading dim ldb) right-hand sides on
 * input (NULL: solve for the assembled FI and return nothing), solutions on
 * output.  The factorisation is kept for later qd_solve calls with the same
 * precision until qd_ghmat8 runs again.  tol bounds the mixed precisions'
 * refinement: it stops once ||b - Ax||inf / (||A||inf ||x||inf + ||b||inf)
 * <= tol; tol <= 0 selects max(1e-14, 2^-53 sqrt(N)), the bound LAPACK's
 * DSGESV stops at.  max_iter <= 0 selects 30 sweeps. */
int qd_solve(qd_handle h, int precision, int nrhs, double* b, int ldb, double tol, int max_iter,
             qd_solve_report* rep);
/* Same with device right-hand sides (overwritten by the solutions). */
int qd_solve_device(qd_handle h, int precision, int nrhs, double* b_dev, long long ldb,
                    double tol, int max_iter, qd_solve_report* rep);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops that solve new right-hand sides against factors
 * already held, code that must not call hipMalloc): qd_solve_device with the
 * factors of an earlier qd_solve / qd_solve_device of the same precision on
 * this handle (QD_ERR_STATE otherwise), and every work vector in one static
 * 64 MiB device workspace per GPU, so the call allocates nothing.  The
 * workspace holds N up to about 100,000 solved one right-hand side at a time,
 * about 36,000 with eight; beyond that the call returns QD_ERR_ARG.  Calls
 * that overlap on one GPU share the workspace and race. */
int qd_solve_device_noalloc(qd_handle h, int precision, int nrhs, double* b_dev, long long ldb,
                            double tol, int max_iter, qd_solve_report* rep);
/* Adjoint solve A^T y = c with the factors of the last qd_solve of this
 * precision (no reference counterpart: the gradient of the solve). */
int qd_solve_transpose_device(qd_handle h, int precision, int nrhs, double* c_dev,
                              long long ldc, double tol, int max_iter, qd_solve_report* rep);

/* reorganize_solution (io_mod.f90:148-183): X (3n) -> FI (3n) displacements,
 * DFI (27, ne) tractions. */
int qd_reorganize_solution(qd_handle h, const double* x, double* fi, double* dfi);

/* compute_internal_points (stress_mod.f90:511-781) for l points.
 * Outputs u_int (3, l), sigma_int (6, l) in xx,yy,zz,xy,yz,zx order,
 * prin (3, l) descending, vm (l); any output may be NULL.
 * counts (nullable, 2 entries): far and near Gauss-point evaluations. */
int qd_internal_points(qd_handle h, int l, const double* cx, const double* cy, const double* cz,
                       const double* fi, const double* dfi, int variant, double* u_int,
                       double* sigma_int, double* prin, double* vm, long long* counts);
/* Same, every array a device pointer (points as (l, 3) row-major xyz). */
int qd_internal_points_device(qd_handle h, int l, const double* pts_dev, const double* fi_dev,
                              const double* dfi_dev, int variant, double* u_dev,
                              double* sigma_dev, double* prin_dev, double* vm_dev,
                              long long* counts);

/* compute_boundary_stresses (stress_mod.f90:366-406): nodal averaged stress
 * (6, n), principal (3, n), von Mises (n); sigma_elem (6, 9, ne) nullable. */
int qd_boundary_stresses(qd_handle h, const double* fi, const double* dfi, double* sigma_elem,
                         double* sigma_node, double* prin, double* vm);

/* ---- text outputs --------------------------------------------------------
 * The reference's result files, record for record in its FORMATs, from host
 * arrays in its layouts and the handle's mesh, material and codes.  Paths
 * are plain file names; NULL skips that file where noted. */
/* The deck echo input_data writes at the top of the output file
 * (io_mod.f90:44-140; the file is created or truncated).  title: the deck's
 * first line; cx, cy, cz the l internal points; ntrack the tracked node. */
int qd_write_input_echo(qd_handle h, const char* path, const char* title, int l, const double* cx,
                        const double* cy, const double* cz, int ntrack);
/* output_data (io_mod.f90:191-228), appended to path: boundary
 * displacements fi (3n), element tractions dfi (27, ne), internal
 * displacements dsol (3l). */
int qd_output_data(qd_handle h, const char* path, const double* fi, const double* dfi, int l, const double* dsol);
/* output_boundary_stresses (stress_mod.f90:412-468): the nodal and
 * element-nodal tables appended to path (nullable) and stresses.dat written
 * to stresses_path (nullable), from qd_boundary_stresses' outputs. */
int qd_output_boundary_stresses(qd_handle h, const char* path, const char* stresses_path,
                                const double* sigma_elem, const double* sigma_node, const double* prin,
                                const double* vm);
/* output_internal_points (stress_mod.f90:892-936): the stress table appended
 * to path (nullable) and internal_points.dat written to points_path
 * (nullable); nothing for l == 0. */
int qd_output_internal_points(qd_handle h, const char* path, const char* points_path, int l, const double* cx,
                              const double* cy, const double* cz, const double* u_int,
                              const double* sigma_int, const double* prin, const double* vm);
/* save_disp_trac (io_mod.f90:230-255): the tracked node's displacement and
 * its traction in its first element, one record (appended when append != 0,
 * else the file is created or truncated). */
int qd_save_disp_trac(qd_handle h, const char* path, int append, int ntrack, const double* fi, const double* dfi);

/* ---- derivatives (no reference counterpart: the reference computes no
 * gradients).  Together with qd_solve_transpose(_device) they give the adjoint
 * gradient of any objective of the internal-point results with respect to
 * the node coordinates and Poisson's ratio. ------------------------------- */
/* New node coordinates for the same connectivity, material and boundary
 * codes (the x, y, z of qd_set_mesh); the geometry, system and factors are
 * recomputed by the next calls. */
int qd_set_coordinates(qd_handle h, const double* x, const double* y, const double* z);
/* Host-buffer form of qd_solve_transpose_device: A^T y = c in place. */
int qd_solve_transpose(qd_handle h, int precision, int nrhs, double* c, int ldc, double tol, int max_iter,
                       qd_solve_report* rep);
/* Derivatives of R = lambda^T (A X - FI) with X held fixed, A and FI being
 * ghmat8's (assembly_mod.f90:53-223) at the current coordinates: gx, gy, gz
 * (n each) = dR/d(x, y, z), *gnu = dR/dnu (mu fixed).  X enters through its
 * reorganised form (qd_reorganize_solution): fi (3n) nodal displacements,
 * dfi (27, ne) element-node tractions.  Outputs are overwritten. */
int qd_ghmat8_vjp(qd_handle h, const double* lambda, const double* fi, const double* dfi, double* gx, double* gy,
                  double* gz, double* gnu);
/* Derivatives of sum_p g_u(:,p) . u_int(:,p) + g_sigma(:,p) . sigma_int(:,p)
 * for compute_internal_points (stress_mod.f90:511-781) at the given fi, dfi:
 * g_fi (3n), g_dfi (27, ne), gx, gy, gz (n), *gnu.  g_u (3, l) and
 * g_sigma (6, l) may be NULL (zero).  Outputs are overwritten. */
int qd_internal_points_vjp(qd_handle h, int l, const double* cx, const double* cy, const double* cz,
                           const double* fi, const double* dfi, int variant, const double* g_u,
                           const double* g_sigma, double* g_fi, double* g_dfi, double* gx, double* gy, double* gz,
                           double* gnu);

/* ---- instrumentation (no reference counterpart: the reference only has
 * cpu_time stopwatches, main.f90:27-71) ------------------------------------
 * With profiling on for a kernel class, every launch of that class is
 * bracketed by HIP events on the handle's stream; qd_profile_read returns
 * the launch count, the summed device time (ms) and the summed algorithmic
 * work (flops for GEMMs, H,G entries for the assembly passes, Gauss-point
 * evaluations for the Somigliana passes). */
enum {
  QD_PROF_LU_TRAILING_GEMM = 0,  /* A22 -= L21 U12 of the outer LU steps */
  QD_PROF_LU_PANEL = 1,          /* panel factorisation (all its launches) */
  QD_PROF_ASSEMBLY_H = 2,        /* H pass of ghmat8 */
  QD_PROF_ASSEMBLY_G = 3,        /* G pass of ghmat8 */
  QD_PROF_SOMIGLIANA_FAR = 4,    /* far-field pass of the internal points */
  QD_PROF_SOMIGLIANA_NEAR = 5,   /* near-field pass of the internal points */
  QD_PROF_KINDS = 6
};
int qd_profile_enable(qd_handle h, unsigned mask /* bit per QD_PROF_* */);
/* The outer trailing update of the FP16 factorisation as an operator (no
 * reference counterpart), on device buffers and the handle's stream:
 *   C[i][j] -= colscale[j / 64] * sum_k A16[i][k] * B16[j][k]
 * A16 (M x K) and B16 (N x K) row-major _Float16 with K a multiple of 32 (pad
 * with zeros), C (M x N) row-major float with leading dimension ldc, colscale
 * ceil(N / 64) floats.  MFMA f16 with FP32 accumulation. */
int qd_fp16_update(qd_handle h, long long M, long long N, int K, const void* A16, const void* B16,
                   const float* colscale, float* C, long long ldc);
/* Device memory held by the library's buffers across all handles in this
 * process: current bytes and the high-water mark since the last reset
 * (reset_peak != 0 resets it to the current value after reading). */
int qd_memory_usage(long long* current_bytes, long long* peak_bytes, int reset_peak);

int qd_profile_read(qd_handle h, int kind, long long* launches, double* ms, double* work);

/* ---- LAPACK drop-in ------------------------------------------------------
 * DGESV with the LAPACK Fortran ABI (the call at solver_mod.f90:28): a gets
 * the LU factors (column-major), ipiv the 1-based row interchanges, b the
 * solutions, info the LAPACK status.  Runs on device 0 (QUADESTH_DEVICE
 * overrides) with FP64 factorisation. */
void dgesv_(const int* n, const int* nrhs, double* a, const int* lda, int* ipiv, double* b,
            const int* ldb, int* info);

/* Device-side DGETRF/DGETRS pair used by dgesv_: a_dev column-major
 * (n x n, leading dim lda), ipiv host (1-based), b_dev column-major. */
int qd_dgesv_device(qd_handle h, int n, int nrhs, double* a_dev, long long lda, int* ipiv,
                    double* b_dev, long long ldb);

#ifdef __cplusplus
}
#endif
#endif
