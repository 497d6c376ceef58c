/*
 * hfo.h -- CPU oracle for the HiFlow3 lmp / linear_solver hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference's CPU path (CPUsimple_CSR_lMatrix / CPUsimple_lVector, the
 * lPreconditioner_* applies and the CG / GMRES / FGMRES / MPIR drivers).  It is
 * the checker that the HIP platform is compared against, and the timed CPU
 * baseline.  Only tests/, the smoke check and bench.py's cpu_baseline leg may
 * load it; the product path never does.
 *
 * Every routine keeps the reference's loop order so that results agree to the
 * last bit with the reference CPU build (compiled without FMA contraction).
 * Each function cites the reference file:line it follows (paths relative to
 * the HiFlow3 2.0 source tree).
 *
 * Parity pins (what each function is checked against) are listed in
 * oracle/README.md.  Every routine has an independent check.
 */
#ifndef HFO_H
#define HFO_H

#ifdef __cplusplus
extern "C" {
#endif

/* IterateControl::State (src/common/iterate_control.h) */
enum { HFO_ITERATE = 0, HFO_SUCCESS_ABS = 1, HFO_SUCCESS_REL = 2,
       HFO_FAIL_DIV = 3, HFO_FAIL_MAXITS = 4 };
/* LinearSolverState (src/linear_solver/preconditioner.h:32-37) */
enum { HFO_SOLVER_SUCCESS = 0, HFO_SOLVER_EXCEEDED = 1, HFO_SOLVER_ERROR = 2 };
/* preconditioner kinds */
enum { HFO_PREC_NONE = 0, HFO_PREC_JACOBI = 1, HFO_PREC_SGS = 2,
       HFO_PREC_MCSGS = 3, HFO_PREC_ILU0 = 4 };

typedef struct {
    int maxits;
    double atol, rtol, dtol;
} hfo_ctrl;

typedef struct {
    int iters;      /* LinearSolver::iter() */
    int status;     /* IterateControl state at exit */
    int ret;        /* LinearSolverState */
    double res;     /* LinearSolver::res() (recursive / estimated residual) */
    double res0;    /* first residual seen by IterateControl */
} hfo_result;

/* Preconditioner description.  For JACOBI `invd` holds 1/a_ii.  For SGS the
 * matrix (row/col/val) is the operator.  For MCSGS the matrix is the colour
 * permuted operator and color_sizes/ncolors describe the colour blocks.  For
 * ILU0 the matrix holds the in-place ILU(0) factors. */
typedef struct {
    int kind;
    const double *invd;
    const int *row, *col;
    const double *val;
    int ncolors;
    const int *color_sizes;
    double omega;
} hfo_prec_d;

typedef struct {
    int kind;
    const float *invd;
    const int *row, *col;
    const float *val;
    int ncolors;
    const int *color_sizes;
    float omega;
} hfo_prec_f;

/* ---------------- CSR construction (integer work) ---------------- */
/* CPU_CSR_lMatrix::TransformFromCOO (lmatrix_csr_cpu.cc:951-1044): bucket by
 * row, then per-row selection sort of the columns.  `data` may be NULL
 * (init_structure).  row_ptr has nrow+1 entries. */
void hfo_csr_from_coo_d(int nrow, int nnz, const int *rows, const int *cols,
                        const double *data, int *row_ptr, int *col, double *val);
/* CPU_CSR_lMatrix::add_values (lmatrix_csr_cpu.cc:1101-1184): dense block of
 * values added at (rows[i], cols[j]) if the entry exists in the pattern. */
void hfo_csr_add_values_d(int nrow, const int *row_ptr, const int *col, double *val,
                          const int *rows, int nr, const int *cols, int nc,
                          const double *values);
/* CPU_CSR_lMatrix::Reorder (lmatrix_csr_cpu.cc:1847-1862 ->
 * lmatrix_coo_cpu.cc:1019-1054): symmetric permutation with perm[new]=old. */
void hfo_csr_reorder_d(int n, const int *row, const int *col, const double *val,
                       const int *perm, int *prow, int *pcol, double *pval);
/* CPU_lVector::Reorder (lvector_cpu.cc:75-93): out[i] = in[perm[i]] */
void hfo_vec_reorder_d(int n, const double *in, const int *perm, double *out);
/* CPU_CSR_lMatrix::Multicoloring (lmatrix_csr_cpu.cc:2175-2291).
 * color_of[i] is the 1-based colour, color_sizes has n entries of room. */
int hfo_multicoloring(int n, const int *row, const int *col, int *ncolors,
                      int *color_of, int *color_sizes, int *perm);
/* CPU_CSR_lMatrix::Levelscheduling (lmatrix_csr_cpu.cc:2295-2391).  Returns
 * the number of levels the algorithm computed; *nlevels_reported receives the
 * value the reference reports (0 when the last level is found in the last
 * sweep; see the function comment). */
int hfo_levelscheduling(int n, const int *row, const int *col,
                        int *nlevels_reported, int *level_sizes, int *perm);
/* lPreconditioner_MultiColoring::dropoffs (lpreconditioner_mc.cc:610-662):
 * zero intra-colour off-diagonals; returns the count. (compress_me is not
 * applied: the zero entries stay in the pattern, which changes no sum.) */
int hfo_mc_dropoffs_d(int n, int *row, int *col, double *val, int ncolors,
                      const int *color_sizes);

/* Threads of the row-parallel loops (SpMV, element-wise updates, the rows of
 * one colour of the multicolour sweeps); 1 (the default) runs everything in
 * one thread.  The results do not depend on it: dot products and the
 * level-ordered sweeps stay sequential in the reference's order. */
void hfo_set_threads(int t);
int hfo_get_threads(void);

/* ---------------- type-generic kernels: _d (double) and _f (float) ---------------- */
#define HFO_DECL(T, S)                                                                  \
void   hfo_spmv_##S(int n, const int *row, const int *col, const T *val,                 \
                    const T *x, T *y);                                                   \
void   hfo_spmv_omp_##S(int n, const int *row, const int *col, const T *val,             \
                        const T *x, T *y);                                               \
void   hfo_spmv_add_##S(int n, const int *row, const int *col, const T *val,             \
                        const T *x, T *y);                                               \
T      hfo_dot_##S(int n, const T *x, const T *y);                                       \
T      hfo_nrm2_##S(int n, const T *x);                                                  \
void   hfo_axpy_##S(int n, T *y, const T *x, T a);                                       \
void   hfo_scaleadd_##S(int n, T *y, T a, const T *x);                                   \
void   hfo_scale_##S(int n, T *y, T a);                                                  \
void   hfo_ewmult_##S(int n, T *y, const T *x);                                          \
void   hfo_invdiag_##S(int n, const int *row, const int *col, const T *val, T *invd);    \
void   hfo_sgs_##S(int n, const int *row, const int *col, const T *val,                  \
                   const T *in, T *out);                                                 \
int    hfo_ilu0_##S(int n, const int *row, const int *col, T *val);                      \
void   hfo_ilu_solve_##S(int n, const int *row, const int *col, const T *lu,             \
                         const T *in, T *out);                                           \
void   hfo_mcsgs_##S(int n, const int *row, const int *col, const T *val,                \
                     int ncolors, const int *color_sizes, T omega,                       \
                     const T *in, T *out);                                               \
void   hfo_prec_apply_##S(int n, const hfo_prec_##S *P, const T *in, T *out);            \
int    hfo_cg_##S(int n, const int *row, const int *col, const T *val, const T *b, T *x, \
                  const hfo_prec_##S *P, const hfo_ctrl *ctrl, hfo_result *res,          \
                  double *hist);                                                         \
int    hfo_gmres_##S(int n, const int *row, const int *col, const T *val, const T *b,    \
                     T *x, const hfo_prec_##S *P, int size_basis, const hfo_ctrl *ctrl,  \
                     hfo_result *res, double *hist);                                     \
int    hfo_fgmres_##S(int n, const int *row, const int *col, const T *val, const T *b,   \
                      T *x, const hfo_prec_##S *P, int size_basis, const hfo_ctrl *ctrl, \
                      hfo_result *res, double *hist);                                    \
void   hfo_swap_diag_front_##S(int n, const int *row, int *col, T *val);                  \
T      hfo_norm1_##S(int n, const T *x);                                                  \
int    hfo_argmax_##S(int n, const T *x);                                                 \
int    hfo_argmin_##S(int n, const T *x);                                                 \
void   hfo_add_values_##S(T *y, const int *idx, int m, const T *v);                       \
void   hfo_add_range_##S(T *y, int s, const T *x, int ss, int m, T w);                    \
void   hfo_rot_##S(int n, T *a, T *b, T c, T s);                                          \
void   hfo_rotm_##S(int n, T *u, T *v, const T *p);                                       \
void   hfo_rotg_##S(T *sa, T *sb, T *sc, T *ss);                                          \
void   hfo_zero_rows_##S(const int *row, const int *col, T *val, const int *idx, int m,   \
                         T alpha);                                                        \
void   hfo_zero_cols_##S(int n, const int *row, const int *col, T *val, const int *idx,   \
                         int m, T alpha);                                                 \
void   hfo_mat_scale_##S(int n, const int *row, const int *col, T *val, T a,              \
                         int offdiag_only);                                               \
void   hfo_pjacobi_##S(int n, const int *row, const int *col, const T *val, const T *in,  \
                       T *out);                                                           \
void   hfo_spmv_nodiag_##S(int n, const int *row, const int *col, const T *val,           \
                           const T *x, T *y);                                             \
int    hfo_jacobi_##S(int n, const int *row, const int *col, const T *val, const T *b,    \
                      T *x, int mode, T w, int inner_iter, int smooth_iter,              \
                      const hfo_ctrl *ctrl, hfo_result *res, double *hist);

HFO_DECL(double, d)
HFO_DECL(float, f)
#undef HFO_DECL

/* CoupledVector::CastInteriorFrom (coupled_vector.cc:755-778): static_cast */
void hfo_cast_d2f(int n, const double *in, float *out);
void hfo_cast_f2d(int n, const float *in, double *out);

/* MPIR<LADescriptorCoupledD, LADescriptorCoupledS>::Solve (mpir.cc:69-304)
 * without pre/post smoothers (their Use flags are off).  The low precision
 * error-correction solver is CG<float> with preconditioner Pl, the high
 * precision one CG<double> with preconditioner Ph. */
typedef struct {
    int max_ech, max_ecl;
    double abs_tol, rel_tol, div_tol;            /* SetMpirParameters */
    hfo_ctrl ecl;                                /* SetLowPrecErrCorSolverParameters */
    hfo_ctrl ech;                                /* SetHighPrecErrCorSolverParameters */
} hfo_mpir_params;

typedef struct {
    int ret;           /* LinearSolverState */
    int iters;         /* accumulated inner iterations */
    int n_ecl, n_ech;  /* error-correction loops taken */
    double res;        /* final true residual norm */
} hfo_mpir_result;

int hfo_mpir(int n, const int *row, const int *col, const double *val_d,
             const float *val_f, const double *b, double *x,
             const hfo_prec_f *Pl, const hfo_prec_d *Ph,
             const hfo_mpir_params *prm, hfo_mpir_result *res, double *hist);

#ifdef __cplusplus
}
#endif
#endif
