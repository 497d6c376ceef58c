/*
 * hflmp.h -- C entry points of the host-side lmp layer (libhflmp.so).
 *
 * Problem generation and the CPU-side preconditioner setup that the
 * reference also performs on the host, plus the lmp-style solver driver that
 * runs the C++ CG / GMRES / FGMRES templates on HIP vectors.  Return value 0
 * on success, negative on invalid input; solver entries return hfh status
 * codes.
 */
#ifndef HFLMP_H
#define HFLMP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Problem kinds: 1 2D Q1 FE Laplacian (nx*ny interior nodes, p0 = coefficient),
 * 2 3D Q1 FE Laplacian (h = 1/(nx+1)), 3 FV diffusion with k = exp(p0 g),
 * 5-pt when nz == 1 else 7-pt, 4 upwind convection-diffusion (p0 = eps,
 * p1 = cell Peclet).  rhs (nullable) is uniform in [-1, 1). */
int hfl_problem_size(int kind, int nx, int ny, int nz, double p0, double p1, unsigned long long seed,
                     int *n, long long *nnz);
int hfl_problem_fill(int kind, int nx, int ny, int nz, double p0, double p1, unsigned long long seed,
                     int *row, int *col, double *val, double *rhs);
/* cell coefficients k = exp(sigma g) and rhs of the FV diffusion problem
 * (kind 3), the values hfl_problem_fill uses; either output may be null */
int hfl_fv_cells(int nx, int ny, int nz, double sigma, unsigned long long seed, double *kcell, double *rhs);
double hfl_uniform_pm1(unsigned long long seed, unsigned long long stream, unsigned long long index);
double hfl_normal01(unsigned long long seed, unsigned long long stream, unsigned long long index);

/* CPU_CSR_lMatrix::Multicoloring (lmp/lmatrix_csr_cpu.cc:2175-2291).
 * color_sizes needs room for n entries. */
int hfl_multicoloring(int n, const int *row, const int *col, int *ncolors, int *color_of,
                      int *color_sizes, int *perm);
/* level schedules of the forward / backward triangular sweeps
 * (CPU_CSR_lMatrix::Levelscheduling, lmp/lmatrix_csr_cpu.cc:2295-2391);
 * level_ptr needs n+1 entries, level_rows n. */
int hfl_levels_lower(int n, const int *row, const int *col, int *nlev, int *level_ptr, int *level_rows);
int hfl_levels_upper(int n, const int *row, const int *col, int *nlev, int *level_ptr, int *level_rows);
/* CPU_CSR_lMatrix::ilu0 (lmp/lmatrix_csr_cpu.cc:1880-1930), in place;
 * returns the number of missing diagonal entries */
int hfl_ilu0(int n, const int *row, const int *col, double *val);
/* CPU_CSR_lMatrix::Reorder (lmp/lmatrix_csr_cpu.cc:1847-1862), perm[new] = old */
int hfl_csr_permute(int n, const int *row, const int *col, const double *val, const int *perm,
                    int *prow, int *pcol, double *pval);

/* ---- MatrixMarket files (CPU_COO_lMatrix::ReadFile / WriteFile,
 * lmatrix_coo_cpu.cc:133-309) and text vectors (CPU_lVector::ReadFile /
 * WriteFile, lvector_cpu.cc:582-660) ----
 * hfl_mm_read: coordinate real/integer, general or symmetric (mirrored
 * entries added) into CSR with ascending columns.  With row == NULL only the
 * sizes are returned (nnz after mirroring); row needs nrow+1 entries, col/val
 * nnz.  hfl_mm_write: "row col value" triples, 1-based, %24.14e.
 * hfl_vec_read: *n in = capacity of values (ignored when values is NULL),
 * out = values read.  hfl_vec_write: std::scientific, 6 digits, as the
 * reference writes. */
int hfl_mm_read(const char *path, int *nrow, int *ncol, long long *nnz, int *row, int *col, double *val);
int hfl_mm_write(const char *path, int nrow, int ncol, const int *row, const int *col, const double *val);
int hfl_vec_read(const char *path, int *n, double *values);
int hfl_vec_write(const char *path, int n, const double *values);

/* ---- the LinearSolver templates on the HIP platform ----
 * CG<LAD> (linear_solver/cg.cc:84-248), GMRES<LAD> (gmres.cc:285-510) and
 * FGMRES<LAD> (fgmres.cc:75-262) with PreconditionerBlockJacobiStand
 * (Jacobi / SGS / ILU(0), preconditioner_bjacobi_standard.cc:45-150) or
 * PreconditionerMultiColoring (SGS, preconditioner_mc.cc:49-194), over
 * single-rank CoupledVector / CoupledMatrix of HIP lVectors / lMatrices.
 * Every Dot returns a host scalar, as through the reference interface. */
enum hfl_solver { HFL_SOLVER_CG = 0, HFL_SOLVER_GMRES = 1, HFL_SOLVER_FGMRES = 2 };
enum hfl_prec { HFL_PREC_NONE = 0, HFL_PREC_JACOBI = 1, HFL_PREC_SGS = 2, HFL_PREC_MCSGS = 3,
                HFL_PREC_ILU0 = 4 };

typedef struct {
    int maxits;
    double atol, rtol, dtol;
} hfl_solve_params;

typedef struct {
    int iters;       /* LinearSolver::iter() */
    int ret;         /* LinearSolverState: 0 success, 1 exceeded, 2 error */
    int state;       /* IterateControl state (MPIR: low-precision loops taken) */
    double res;      /* LinearSolver::res() */
    double *hist;    /* optional: residual history (caller buffer) */
    int hist_cap;    /* capacity of hist */
    int nhist;       /* entries written */
} hfl_solve_result;

const char *hfl_last_error(void);
/* Solves A x = b (x: start vector in, solution out) on device 0.  single != 0
 * runs the FP32 descriptor.  For HFL_PREC_MCSGS the system is coloured and
 * renumbered first and x is returned in the original numbering. */
int hfl_lmp_solve(int solver, int precond, int size_basis, double omega, int single, int n,
                  const int *row, const int *col, const double *val, const double *b, double *x,
                  const hfl_solve_params *params, hfl_solve_result *result);
/* MPIR<LADescriptorHipD, LADescriptorHipS> (mpir.cc:69-327) with a CG<float>
 * error-correction solver (Jacobi or no preconditioner). */
int hfl_lmp_mpir(int precond_low, int n, const int *row, const int *col, const double *val,
                 const double *b, double *x, int max_ecl, double abs_tol, double rel_tol, double div_tol,
                 const hfl_solve_params *ecl, hfl_solve_result *result);

#ifdef __cplusplus
}
void hflmp_set_error(const char *msg);
#endif
#endif
