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

#ifdef __cplusplus
}
#endif
#endif
