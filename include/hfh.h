/*
 * hfh.h -- C ABI of the MI355X (gfx950) HIP platform for HiFlow3's lmp layer.
 *
 * Plain pointers and sizes only.  Every function returns an int status
 * (HFH_OK == 0, negative codes on failure) and never exits the process;
 * hfh_last_error() returns the message of the calling thread's last failure.
 * All device work of a context is ordered on the context's single HIP stream;
 * calls that return a host value (dot, nrm2, downloads, solves) synchronise.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the HiFlow3 2.0 source tree, src/).
 */
#ifndef HFH_H
#define HFH_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hfh_ctx hfh_ctx;
typedef struct hfh_vec hfh_vec;
typedef struct hfh_csr hfh_csr;
typedef struct hfh_prec hfh_prec;

enum hfh_status {
    HFH_OK = 0,
    HFH_ERR_INVALID = -1,       /* bad argument or inconsistent sizes */
    HFH_ERR_NOMEM = -2,         /* device allocation failed */
    HFH_ERR_HIP = -3,           /* HIP runtime error */
    HFH_ERR_ZERO_PIVOT = -4,    /* zero / missing diagonal entry */
    HFH_ERR_NOT_CONVERGED = -5, /* solver stopped on maxits or divergence */
    HFH_ERR_BUDGET = -6,        /* serving admission over the HBM budget */
    HFH_ERR_UNSUPPORTED = -7    /* dtype / option combination not provided */
};

enum hfh_dtype { HFH_F64 = 0, HFH_F32 = 1 };

/* IterateControl states (common/iterate_control.h) */
enum hfh_iter_state {
    HFH_ITERATE = 0, HFH_SUCCESS_ABS = 1, HFH_SUCCESS_REL = 2,
    HFH_FAIL_DIV = 3, HFH_FAIL_MAXITS = 4
};

enum hfh_prec_kind {
    HFH_PREC_NONE = 0, HFH_PREC_JACOBI = 1, HFH_PREC_SGS = 2,
    HFH_PREC_MCSGS = 3, HFH_PREC_ILU0 = 4
};

const char *hfh_last_error(void);
const char *hfh_version(void);

/* ---- context: platform init (lmp/platform_management.cc:135-187) ---- */
int hfh_device_count(int *count);
/* The drivers replay captured hipGraphs of their iteration chunks; with the
 * environment variable HFH_GRAPHS=0 (read once per process) they launch the
 * same kernels directly instead (same results), e.g. for tracing tools. */
int hfh_ctx_create(int device, hfh_ctx **ctx);
int hfh_ctx_destroy(hfh_ctx *ctx);
int hfh_ctx_sync(hfh_ctx *ctx);
/* device properties used by the launch heuristics */
int hfh_ctx_info(hfh_ctx *ctx, int *num_cu, long long *hbm_bytes, char *name, int name_len);
/* per-kernel-class HIP-event timing; class: 0 = SpMV, 1 = PCG step, 2 = precond */
int hfh_ctx_set_timing(hfh_ctx *ctx, int on);
int hfh_ctx_timing(hfh_ctx *ctx, int kclass, long long *launches, double *total_ms);
int hfh_ctx_timing_reset(hfh_ctx *ctx);

/* ---- vectors: lVector<T> (lmp/lvector.h:39-336) ---- */
/* lVector::Init (lvector.h:199) + CloneWithoutContent (lvector.cc:52-64): zero filled */
int hfh_vec_create(hfh_ctx *ctx, int n, int dtype, hfh_vec **v);
int hfh_vec_destroy(hfh_vec *v);
int hfh_vec_size(const hfh_vec *v, int *n);
int hfh_vec_dtype(const hfh_vec *v, int *dtype);
/* host <-> device: SetBlockValues / GetBlockValues (lvector.h:112-126),
 * CopyFrom / CopyTo across platforms (lvector_gpu.cc:389-435) */
int hfh_vec_upload(hfh_vec *v, const void *host, int offset, int count);
int hfh_vec_download(const hfh_vec *v, void *host, int offset, int count);
/* SetValues / GetValues with a host index set (lvector.h:93-110) */
int hfh_vec_set_values(hfh_vec *v, const int *index, int count, const void *values);
int hfh_vec_get_values(const hfh_vec *v, const int *index, int count, void *values);
/* Zeros (lvector.h:206) */
int hfh_vec_zeros(hfh_vec *v);
/* CopyFrom (lvector.h:222), same platform */
int hfh_vec_copy(hfh_vec *dst, const hfh_vec *src);
/* partial_replace_subvector (lvector.h:177-181):
 * dst[dst_off + i] = src[src_off + i], i < count */
int hfh_vec_copy_range(hfh_vec *dst, int dst_off, const hfh_vec *src, int src_off, int count);
/* CastFrom (lvector.h:231-232): f64 <-> f32 with round-to-nearest */
int hfh_vec_cast(hfh_vec *dst, const hfh_vec *src);
/* Reorder (lvector.h:216): dst[i] = src[perm[i]], perm on the host */
int hfh_vec_reorder(hfh_vec *dst, const hfh_vec *src, const int *perm);

/* ---- BLAS-1 (CPUsimple_lvector.cc:142-435; lvector.h:262-291) ---- */
/* Dot: blocking, returns the host scalar (CoupledVector::Dot, coupled_vector.cc:330-344) */
int hfh_dot(const hfh_vec *x, const hfh_vec *y, double *out);
/* device address of a vector's data (for interop with other device code;
 * the handle keeps ownership) */
int hfh_vec_device_ptr(const hfh_vec *v, void **ptr);
/* Dot without a host wait: the FP64 result is stored at dev_out (device
 * memory) in stream order; a later hfh call or hfh_ctx_sync sees it
 * (the asynchronous form of lVector::Dot, lvector.h:274) */
int hfh_dot_dev(const hfh_vec *x, const hfh_vec *y, double *dev_out);
/* Norm2 = sqrt(Dot(self)) (CPUsimple_lvector.cc:142-152) */
int hfh_nrm2(const hfh_vec *x, double *out);
/* Axpy: y += a*x (CPUsimple_lvector.cc:234-287) */
int hfh_axpy(hfh_vec *y, const hfh_vec *x, double a);
/* ScaleAdd: y = a*y + x (CPUsimple_lvector.cc:310-365) */
int hfh_scaleadd(hfh_vec *y, double a, const hfh_vec *x);
/* Scale: y *= a (CPUsimple_lvector.cc:388-435) */
int hfh_scale(hfh_vec *y, double a);
/* ElementWiseMult: y .*= x (lvector_cpu.cc:571-578) */
int hfh_ewmult(hfh_vec *y, const hfh_vec *x);
/* Norm1: sum |x_i| (CPUsimple_lvector.cc:95-140); FP64 partials per
 * workgroup summed in a fixed order (the result differs from the reference's
 * unroll-5 order in the last bits, as Dot does).  HFH_ERR_INVALID when empty. */
int hfh_norm1(const hfh_vec *x, double *out);
/* ArgMax / ArgMin (CPUsimple_lvector.cc:51-93): the first index of the
 * largest / smallest |x_i|, with the reference loop's start at |x_0|: NaN
 * entries are never chosen, and a NaN x_0 answers 0.  HFH_ERR_INVALID when
 * empty. */
int hfh_iamax(const hfh_vec *x, int *index);
int hfh_iamin(const hfh_vec *x, int *index);
/* NormMax = |x[ArgMax]| (CPUsimple_lvector.cc:154-157) */
int hfh_normmax(const hfh_vec *x, double *out);
/* add_values (lvector_cpu.cc:123-135): v[index[i]] += values[i], a repeated
 * index adding its values in their order (add_value is count = 1) */
int hfh_vec_add_values(hfh_vec *v, const int *index, int count, const void *values);
/* partial_add_subvector (lvector_cpu.cc:504-548):
 * dst[dst_off + i] += weight * src[src_off + i], i < count */
int hfh_vec_add_range(hfh_vec *dst, int dst_off, const hfh_vec *src, int src_off, int count, double weight);
/* set_indexset / get_indexset / get_indexset_size / GetIndexedValues /
 * SetIndexedValues (lvector_cpu.cc:217-297): the index set lives on the
 * device with the vector; the values move between the host and the indexed
 * entries (a gather / scatter kernel); a repeated index takes its last value
 * on SetIndexedValues, as the sequential loop does.  get_indexset: index may
 * be NULL to query the count. */
int hfh_vec_set_indexset(hfh_vec *v, const int *index, int count);
int hfh_vec_get_indexset(const hfh_vec *v, int *index, int *count);
int hfh_vec_get_indexed_values(const hfh_vec *v, void *values);
int hfh_vec_set_indexed_values(hfh_vec *v, const void *values);
/* Rot (CPUsimple_lvector.cc:448-466): x_i' = c x_i + s y_i, y_i' = c y_i - s x_i
 * (x is the vector Rot is called on, y its argument) */
int hfh_rot(hfh_vec *x, hfh_vec *y, double c, double s);
/* Rotm (CPUsimple_lvector.cc:540-600), param = {flag, h11, h21, h12, h22}:
 * u is the vector Rotm is called on, v its argument, with the reference's
 * roles: flag -1: v' = h11 v + h12 u, u' = h21 v + h22 u; flag 0: the same
 * with h11 = h22 = 1; flag 1: v' = h11 v + u, u' = h22 u - v; any other flag
 * (-2): no change */
int hfh_rotm(hfh_vec *u, hfh_vec *v, const double *param);

/* ---- CSR matrices: CSR_lMatrixType<T> (lmp/lmatrix_formats.h:40-46) ---- */
/* lMatrix::Init (lmatrix.h:107) */
int hfh_csr_create(hfh_ctx *ctx, int nrow, int ncol, int nnz, int dtype, hfh_csr **A);
int hfh_csr_destroy(hfh_csr *A);
int hfh_csr_dims(const hfh_csr *A, int *nrow, int *ncol, int *nnz, int *dtype);
/* CopyStructureFrom + CopyFrom from a host CSR (lmatrix_csr_gpu.cc:140-192).
 * Validates row[0]==0, monotone rows, row[nrow]==nnz, columns in range and
 * strictly ascending within each row (lmatrix_csr_cpu.cc:1017-1044). */
int hfh_csr_upload(hfh_csr *A, const int *row, const int *col, const void *val);
/* CopyFrom with an unchanged structure: values only */
int hfh_csr_upload_values(hfh_csr *A, const void *val);
int hfh_csr_download(const hfh_csr *A, int *row, int *col, void *val);
/* CastFrom (lmatrix.h:297-298): dst takes src's structure, values converted.
 * SpMV layout: dst's requested one (hfh_csr_set_spmv), else src's requested
 * one, else the automatic choice for dst's value type. */
int hfh_csr_cast(hfh_csr *dst, const hfh_csr *src);
/* VectorMult: y = A x (lmatrix.h:331; CPUsimple_csr_lmatrix.cc:75-105) */
int hfh_spmv(const hfh_csr *A, const hfh_vec *x, hfh_vec *y);
/* VectorMultAdd: y += A x (lmatrix.h:336; lmatrix_csr_cpu.cc:1354-1375) */
int hfh_spmv_add(const hfh_csr *A, const hfh_vec *x, hfh_vec *y);
/* SpMV kernel choice (no reference counterpart; the reference selects its
 * SpMV by platform/implementation enums, la_global.h:40-50).
 * HFH_SPMV_STREAM: row-block LDS streaming, param1 = nonzeros per block
 * (512/1024/2048/4096, 0 = 2048), bit-identical to the CPU row sums;
 * HFH_SPMV_VECTOR: param1 lanes per row (4..64, 0 = from the mean row
 * length); HFH_SPMV_WAVE: per-wavefront tiles of param1 nonzeros (256/512/
 * 1024) and at most 63 rows, bit-identical like STREAM; HFH_SPMV_SELL: a
 * SELL-64 copy (slices of 64 rows, column-major, built on the device from
 * the CSR arrays), lane = row, param1 column slots in flight (2/4/8/16),
 * bit-identical like STREAM.
 * param2: workgroups (0 = automatic).  algo 0 = automatic.
 * Takes effect at once on an uploaded matrix and on later uploads. */
enum hfh_spmv_algo { HFH_SPMV_AUTO = 0, HFH_SPMV_STREAM = 1, HFH_SPMV_VECTOR = 2, HFH_SPMV_WAVE = 3,
                     HFH_SPMV_SELL = 4, HFH_SPMV_SELL_DICT = 5, HFH_SPMV_SELL_PAIRS = 6,
                     HFH_SPMV_SELL_LDS = 7 };
/* HFH_SPMV_SELL_LDS: the slot-pair layout with x staged in LDS.  For every
 * slice the distinct 16-element chunks of x its columns touch are listed; a wavefront loads its slice's chunks once with 16-byte loads and
 * each slot holds a 16-bit index into that LDS image (2 index bytes per entry
 * instead of 4, and one coalesced load per chunk instead of one gather per
 * entry).  Bit-identical like SELL.  The automatic choice for rows of 16 or
 * more entries on average where SELL-64 padding stays small and every
 * slice's window fits.  A requested SELL_LDS on a matrix with a slice whose
 * chunks exceed 1024 elements of x gets HFH_SPMV_SELL_PAIRS instead (the
 * automatic choice falls back to its SELL / pairs rule).  param1 unused. */
/* HFH_SPMV_SELL_PAIRS: the SELL-64 copy with the two consecutive column slots
 * of a row stored side by side (16-byte value loads per lane for FP64); param1:
 * slot pairs in flight (1 or 2).  Bit-identical like SELL.  The automatic
 * choice for FP64 matrices with long rows (mean length >= 16) whose x windows
 * do not fit the LDS-staged layout. */
/* HFH_SPMV_SELL_DICT: the SELL-64 copy with its columns stored as one-byte
 * codes into a table of the matrix's distinct column offsets (col - row),
 * four codes of consecutive slots per 32-bit word: 1 index byte per entry
 * instead of 4 (stencil matrices).  param1: code words in flight (1 or 2).
 * Bit-identical like SELL; a matrix with more than 256 distinct offsets gets
 * HFH_SPMV_SELL instead. */
int hfh_csr_set_spmv(hfh_csr *A, int algo, int param1, int param2);
/* the SpMV layout in use (after the automatic choice or a fallback) */
int hfh_csr_get_spmv(const hfh_csr *A, int *algo, int *param1);
/* extract_invdiagelements over all rows (lmatrix_csr_cpu.cc:2414-2432) */
int hfh_csr_invdiag(const hfh_csr *A, hfh_vec *invd);
/* extract_diagelements / extract_invdiagelements over rows [start, end)
 * (lmatrix_csr_cpu.cc:2395-2432): out[i] = a_ii or 1 / a_ii, i = start + k;
 * 0 for a row without a stored diagonal */
int hfh_csr_diag(const hfh_csr *A, int start, int end, int invert, hfh_vec *out);
/* Zeros (lmatrix_csr_cpu.cc:100-104): every value 0, structure kept */
int hfh_csr_zeros(hfh_csr *A);
/* Scale / ScaleOffdiag (lmatrix_csr_cpu.cc:206-225): values *= alpha (only
 * the entries off the diagonal when offdiag_only) */
int hfh_csr_scale(hfh_csr *A, double alpha, int offdiag_only);
/* ZeroRows (lmatrix_csr_cpu.cc:1047-1068): the listed rows' entries 0, their
 * diagonal entries alpha; ZeroCols (lmatrix_csr_cpu.cc:1071-1098): the same
 * for the listed columns (count > 0).  Indices are checked against the
 * matrix's rows / columns. */
int hfh_csr_zero_rows(hfh_csr *A, const int *rows, int count, double alpha);
int hfh_csr_zero_cols(hfh_csr *A, const int *cols, int count, double alpha);
/* Pjacobi (lmatrix_csr_cpu.cc:920-948): out_i = in_i / a_ii; rows without a
 * stored diagonal leave out_i unchanged */
int hfh_csr_pjacobi(const hfh_csr *A, const hfh_vec *in, hfh_vec *out);

/* ---- preconditioners: lPreconditioner<T> (lmp/lpreconditioner.h:35-125) ---- */
/* lPreconditioner_Jacobi::Build (lpreconditioner.cc:127-148) */
int hfh_jacobi_create(const hfh_csr *A, hfh_prec **P);
/* lPreconditioner_SymmetricGaussSeidel (lpreconditioner.cc:221-230 ->
 * lmatrix_csr_cpu.cc:542-616), natural order, solved level by level.
 * level_ptr[nlev+1], level_rows[nrow]: forward levels (Levelscheduling,
 * lmatrix_csr_cpu.cc:2295-2391); the backward sweep uses its own levels
 * computed from the upper triangle on the host. */
int hfh_sgs_create(const hfh_csr *A, int nlev, const int *level_ptr, const int *level_rows,
                   int nlev_up, const int *up_level_ptr, const int *up_level_rows,
                   hfh_prec **P);
/* lPreconditioner_MultiColoring_SymmetricGaussSeidel (lpreconditioner_mc_sgs.cc,
 * lpreconditioner_mc.cc:157-770) on the colour-permuted matrix A.  Colour c
 * holds rows [sum(color_sizes[<c]), +color_sizes[c]).  Intra-colour
 * off-diagonals are ignored (dropoffs). */
int hfh_mcsgs_create(const hfh_csr *A, int ncolors, const int *color_sizes, double omega,
                     hfh_prec **P);
/* lPreconditioner_ILUp with p = 0 (lpreconditioner.cc:388-425): LU holds the
 * ILU(0) factors (CPU_CSR_lMatrix::ilu0, lmatrix_csr_cpu.cc:1880-1930, computed
 * on the host); the apply is ilu_solve (lmatrix_csr_cpu.cc:2108-2171) level by
 * level, with the forward and backward level sets given as for SGS. */
int hfh_ilu0_create(const hfh_csr *LU, int nlev, const int *level_ptr, const int *level_rows,
                    int nlev_up, const int *up_level_ptr, const int *up_level_rows,
                    hfh_prec **P);
/* CPU_CSR_lMatrix::ilu0 (lmatrix_csr_cpu.cc:1880-1930) on the device, in
 * place: the rows of each level of the lower level schedule (as for
 * hfh_sgs_create) are factored in parallel with the same per-row arithmetic,
 * so the factors are bit-identical to the host loop.  *missing: lower entries
 * skipped for lack of a stored diagonal in their row. */
int hfh_csr_ilu0(hfh_csr *A, int nlev, const int *level_ptr, const int *level_rows, int *missing);
/* ---- preconditioner setup on the device ----
 * The same results as the host analyses, computed from the device pattern
 * (a fixed-point iteration over all rows, then a stable sort); outputs are
 * host arrays in the formats the create calls above take.
 * CPU_CSR_lMatrix::Levelscheduling (lmatrix_csr_cpu.cc:2295-2391): upper = 0
 * orders the rows for the forward sweep (dependencies on lower columns),
 * upper = 1 for the backward sweep.  level_ptr: n + 1 entries, level_rows: n. */
int hfh_csr_levels(const hfh_csr *A, int upper, int *nlev, int *level_ptr, int *level_rows);
/* CPU_CSR_lMatrix::Multicoloring (lmatrix_csr_cpu.cc:2175-2291): first-fit
 * colours (1-based, color_of optional, n entries), the count of each colour
 * (color_sizes: room for n entries) and perm[new] = old, rows stable by
 * colour.  HFH_ERR_INVALID for a pattern without off-diagonal entries. */
int hfh_csr_multicoloring(const hfh_csr *A, int *ncolors, int *color_of, int *color_sizes, int *perm);
/* CPU_CSR_lMatrix::Reorder (lmatrix_csr_cpu.cc:1847-1862): P = A permuted
 * symmetrically, row i of P = row perm[i] of A with columns renumbered and
 * ascending.  P: created with A's dims, nnz and dtype; perm (host) is checked
 * to be a permutation. */
int hfh_csr_permute(const hfh_csr *A, const int *perm, hfh_csr *P);

/* ---- the seeded benchmark systems assembled on the device ----
 * No reference counterpart: the reference assembles on the host
 * (CSR_lMatrix::init_structure + add_values, lmatrix_csr_cpu.cc:1101-1184).
 * The rows of hfl_problem_fill (kinds 1 2D Q1, 2 3D Q1, 3 FV diffusion,
 * 4 upwind convection-diffusion; same grid and parameters) are built on the
 * device, bit-identical to the host generator; rhs (optional, device, size n,
 * A's dtype) receives its seeded right-hand side, also bit-identical.  A is
 * created with the n and nnz hfl_problem_size reports; kcell: the n FV cell
 * coefficients from hfl_fv_cells (kind 3 only, host), else NULL. */
int hfh_problem_assemble(hfh_csr *A, int kind, int nx, int ny, int nz, double p0, double p1,
                         unsigned long long seed, const double *kcell, hfh_vec *rhs);
/* ApplylPreconditioner (lpreconditioner.h): out = M^{-1} in */
int hfh_prec_apply(hfh_prec *P, const hfh_vec *in, hfh_vec *out);
int hfh_prec_kind(const hfh_prec *P, int *kind);
int hfh_prec_destroy(hfh_prec *P);

/* Schedule of the SGS / ILU(0) triangular sweeps (no reference counterpart;
 * the reference sweeps row by row on the CPU).  HFH_SWEEP_SYNCFREE (the
 * default): one launch per sweep; waves take work in dependency order from
 * a counter and each row waits for the earlier results it needs (bounded
 * waits).  Where the pattern is that of a natural-order structured grid
 * (found at create time), a wave's work is a geometric block of up to 64
 * rows whose own dependencies resolve between lanes, else a slice of 64 rows
 * of one level.  HFH_SWEEP_SLICES: sync-free by level slices only.
 * HFH_SWEEP_LEVELS: one launch per level (runs of small levels merged).  All
 * give bit-identical results.  waves_per_cu: waves of a sync-free sweep per
 * compute unit (1..32); 0 (the default) sizes it to about three levels'
 * work, between 1 and 8 per compute unit. */
enum { HFH_SWEEP_SYNCFREE = 0, HFH_SWEEP_LEVELS = 1, HFH_SWEEP_SLICES = 2 };
int hfh_prec_set_sweep(hfh_prec *P, int mode, int waves_per_cu);
/* HFH_ERR_HIP when a sync-free sweep of P gave up waiting since the last
 * call (its results are then NaN); clears the condition.  Synchronises the
 * context's stream.  The drivers check it at the end of each solve. */
int hfh_prec_sweep_status(hfh_prec *P);
/* Shape of the blocked sync-free sweeps (no reference counterpart): blocks
 * and block levels of the forward and backward sweeps; 0 blocks where the
 * pattern was not recognised as a structured grid (slices then run). */
int hfh_prec_sweep_blocks(const hfh_prec *P, int *fw_blocks, int *fw_levels, int *bw_blocks, int *bw_levels);

/* ---- device-resident Krylov drivers ---- */
typedef struct {
    int maxits;
    double atol, rtol, dtol;   /* IterateControl (iterate_control.cc:84-110) */
} hfh_control;

typedef struct {
    int iters;       /* LinearSolver::iter() */
    int state;       /* hfh_iter_state at exit */
    double res;      /* last residual given to IterateControl */
    double res0;     /* first residual */
    double time_ms;  /* device time of the solve (HIP events) */
} hfh_solve_info;

/* CG<LAD>::SolvePrecond / SolveNoPrecond (linear_solver/cg.cc:84-248) with
 * the scalars kept on the device and the loop launched in chunks; same
 * iterates and stopping test as the reference.  P may be NULL (no
 * preconditioning), Jacobi or MC-SGS.  Returns HFH_OK on convergence,
 * HFH_ERR_NOT_CONVERGED on maxits/divergence (info is filled either way).
 * An empty system (n = 0) returns at once: 0 iterations, residual 0,
 * state HFH_SUCCESS_ABS (the same holds for GMRES and defect correction).
 * hist (optional, host, maxits+1 doubles) receives the residual history. */
int hfh_pcg_solve(const hfh_csr *A, hfh_prec *P, const hfh_vec *b, hfh_vec *x,
                  const hfh_control *ctl, hfh_solve_info *info, double *hist);

/* hfh_pcg_solve without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The work
 * vectors r, p, z, Ap, the partial sums, the solver state and the history
 * live in a static device array of 64 MiB that each device the library runs
 * on holds until the process ends: 4 n elements plus about 80 KB (and
 * maxits + 2 doubles with hist), so FP64 systems up to about 2.09 M rows,
 * FP32 up to about 4.19 M; larger solves return HFH_ERR_INVALID.  P must be
 * NULL, HFH_PREC_NONE or Jacobi (else HFH_ERR_UNSUPPORTED).  Same iterates,
 * stopping test and results as hfh_pcg_solve; the loop runs as direct
 * launches (no graph) and info->time_ms is host wall time.  Every call on
 * one device shares the array: calls that overlap on one device race. */
int hfh_pcg_solve_noalloc(const hfh_csr *A, hfh_prec *P, const hfh_vec *b, hfh_vec *x,
                          const hfh_control *ctl, hfh_solve_info *info, double *hist);

/* GMRES<LAD>::SolveRight (gmres.cc:285-510) / FGMRES<LAD>::SolveRight
 * (fgmres.cc:75-262) with right preconditioning, modified Gram-Schmidt and
 * Givens rotations; the Hessenberg scalars live on the device.  flexible
 * selects FGMRES.  P may be NULL.  For n above 1024 x (compute units) and up
 * to 8 times that, the MGS sweep of each Arnoldi column runs as one
 * cooperative launch (HFH_ERR_HIP if one of its bounded waits gave up). */
int hfh_gmres_solve(const hfh_csr *A, hfh_prec *P, int size_basis, int flexible,
                    const hfh_vec *b, hfh_vec *x, const hfh_control *ctl,
                    hfh_solve_info *info, double *hist);

/* ---- stationary Jacobi solver / smoother (linear_solver/jacobi.cc) ---- */
/* CSR_lMatrix::SwapDiagElementsToRowFront (lmatrix_csr_cpu.cc:477-495), as
 * Jacobi<LAD>::Prepare (jacobi.cc:44-67) applies it to the operator: each
 * row's diagonal entry trades places with the row's first entry, in place on
 * the device.  The SpMV of the matrix then sums rows in the new stored order.
 * Afterwards the columns are no longer ascending, and the SGS / MC-SGS /
 * ILU(0) setups refuse the matrix (HFH_ERR_INVALID). */
int hfh_csr_swap_diag_front(hfh_csr *A);
/* 1 when every row's columns are ascending (as uploaded), 0 after the swap. */
int hfh_csr_is_sorted(const hfh_csr *A, int *sorted);
/* CPU_CSR_lMatrix::VectorMultNoDiag (lmatrix_csr_cpu.cc:521-538):
 * y_i = sum over j != i of a_ij x_j, in stored order. */
int hfh_spmv_nodiag(const hfh_csr *A, const hfh_vec *x, hfh_vec *y);
/* Jacobi<LAD>::Solve (jacobi.cc:100-285) on one rank.  mode 0 SolveNormal,
 * 1 SolveDamped, 2 SmoothNormal, 3 SmoothDamped (SetSolveMode); w the
 * damping parameter (SetDampingParameter, > 0 for the damped modes);
 * inner_iter steps between residual checks (SetInnerIter); smooth_iter
 * blocks of inner_iter steps in the smoother modes (SetNumIter).  A is the
 * operator after hfh_csr_swap_diag_front, as Prepare leaves it.  The solve
 * modes test ||b - A x|| with ctl (IterateControl, iteration count = steps)
 * and return HFH_ERR_NOT_CONVERGED on maxits / divergence; the smoother modes
 * ignore ctl (may be NULL) and return HFH_OK.  hist (optional, solve modes,
 * maxits/inner_iter + 2 doubles) receives every checked residual. */
int hfh_jacobi_solve(const hfh_csr *A, const hfh_vec *b, hfh_vec *x, int mode, double w,
                     int inner_iter, int smooth_iter, const hfh_control *ctl,
                     hfh_solve_info *info, double *hist);

/* Mixed-precision defect correction (MPIR<D,S>::Solve, mpir.cc:69-327):
 * FP64 true residual, FP64 update and FP64 outer test on the device; FP32
 * matrix values and FP32 inner PCG with preconditioner Pl (Jacobi or MC-SGS
 * built on Al).  The outer loop runs while the true residual is > outer.atol,
 * > outer.rtol * res0 and < outer.dtol, and while a correction is allowed
 * (fewer than outer.maxits so far, residual > inner.atol, > inner.rtol * res0,
 * < inner.dtol).  res0 is reset to the residual just computed after every
 * correction (the reference's res0_ reset, mpir.cc:275), so with rtol < 1
 * the relative tests compare the residual with itself: only atol, dtol and
 * maxits end the loop, and HFH_SUCCESS_REL is not returned.  inner: control
 * of each FP32 inner solve; its tolerances are rounded to float first, as the
 * reference stores them (DataType_low members, mpir.h:225-240).  x and b must
 * not alias. */
typedef struct {
    int outer_iters;     /* corrections applied */
    int inner_iters;     /* accumulated inner iterations */
    int state;           /* HFH_SUCCESS_ABS / _REL, or HFH_FAIL_* */
    double res;          /* final true FP64 residual norm */
    double res0;         /* initial true residual norm */
    double time_ms;
} hfh_dc_info;

int hfh_dc_solve(const hfh_csr *Ah, const hfh_csr *Al, hfh_prec *Pl, const hfh_vec *b,
                 hfh_vec *x, const hfh_control *outer, const hfh_control *inner,
                 hfh_dc_info *info, double *hist);

/* Options of the defect correction.  keep_direction = 0 is hfh_dc_solve:
 * every correction starts a fresh inner CG (MPIR, mpir.cc:160-200).
 * keep_direction = 1: the inner CG keeps its search direction across
 * corrections.  At each FP64 correction the FP32 residual is replaced by the
 * cast true residual, and the next direction is p = z + (r.z / r_k.z_k) p
 * instead of p = z: the reliable-update CG of Sleijpen & van der Vorst (1996),
 * as used for mixed-precision Krylov solvers by Clark et al. (2010).  No
 * reference counterpart; the inner control's rtol then sets how often the
 * FP64 residual is recomputed. */
typedef struct {
    int keep_direction;
} hfh_dc_options;
int hfh_dc_solve_ex(const hfh_csr *Ah, const hfh_csr *Al, hfh_prec *Pl, const hfh_vec *b,
                    hfh_vec *x, const hfh_control *outer, const hfh_control *inner,
                    const hfh_dc_options *opt, hfh_dc_info *info, double *hist);

/* ---- serving: many independent SPD systems solved together ----
 * No reference counterpart: the engine runs, for every admitted system, the
 * defect correction of MPIR<D,S>::Solve (mpir.cc:69-327) with an FP32 inner
 * Jacobi-preconditioned CG (lpreconditioner.cc:104-166) in the
 * Chronopoulos-Gear form of cg.cc:84-248 (the same iterates in exact
 * arithmetic, one reduction per step) and FP64 true residual, update and
 * stop test, for all systems in the same batched launches (a hipGraph of
 * `chunk` steps, three launches each).  Systems are admitted while
 * others run, as slots and the HBM budget allow, and leave when collected.
 * A system's result does not depend on what else is in the batch: its
 * arithmetic and summation order are fixed by its own size.
 * Outer stop: true residual <= max(atol, rtol * initial true residual);
 * inner stop: recursive FP32 residual <= inner_rtol * its start, or
 * inner_maxits.  Calls on one engine must come from one thread. */
typedef struct hfh_batch hfh_batch;
typedef struct {
    int max_systems;       /* resident systems (slots), 1..65536 */
    long long hbm_budget;  /* bytes of the device arena holding admitted systems */
    int chunk;             /* iterations per hfh_batch_step (graph replay), 1..256 */
    double rtol, atol;     /* outer stop */
    int outer_maxits;      /* corrections per system */
    double inner_rtol;     /* inner stop */
    int inner_maxits;      /* iterations per inner solve */
    /* 1..4: systems of at most this many tiles of 1024 rows run 4 * `chunk`
     * iterations per step in one persistent workgroup each (kb_small, on a
     * second stream), beside the batched tile launches of the larger ones;
     * 0 (the default): every system in the tile launches.  Both paths give
     * the same bits. */
    int small_tiles;
} hfh_batch_params;

typedef struct {
    int outer_iters;       /* corrections applied */
    int inner_iters;       /* inner iterations, all corrections */
    int state;             /* HFH_SUCCESS_ABS / _REL or HFH_FAIL_* */
    double res;            /* final true FP64 residual norm */
    double res0;           /* initial true residual norm */
} hfh_batch_result;

int hfh_batch_create(hfh_ctx *ctx, const hfh_batch_params *p, hfh_batch **B);
int hfh_batch_destroy(hfh_batch *B);
/* bytes of the budget a request would take */
int hfh_batch_bytes_csr(int n, const int *row, long long *bytes);
int hfh_batch_bytes_fv(int nx, int ny, int nz, long long *bytes);
/* Admit an SPD CSR system (columns ascending, nonzero diagonal; x0 nullable
 * = zero start).  Returns HFH_ERR_BUDGET when no slot or not enough budget is
 * free now (retry after hfh_batch_step / hfh_batch_collect), HFH_ERR_INVALID
 * when the request can never fit.  Host arrays are copied before return. */
int hfh_batch_submit_csr(hfh_batch *B, int n, const int *row, const int *col, const double *val,
                         const double *b, const double *x0, int *id);
/* Admit the cell-centred FV diffusion system of an nx*ny*nz grid (5-pt when
 * nz == 1, else 7-pt) from its cell coefficients k > 0, assembled on the
 * device with the entries hfl_problem_fill (kind 3) produces. */
int hfh_batch_submit_fv(hfh_batch *B, int nx, int ny, int nz, const double *kcell, const double *b,
                        const double *x0, int *id);
/* Run `chunk` iterations of every admitted system; *active: systems still
 * running, *finished: systems that finished in this step. */
int hfh_batch_step(hfh_batch *B, int *active, int *finished);
/* ids of finished, not yet collected systems (up to cap); *count = all of them */
int hfh_batch_finished(hfh_batch *B, int *ids, int cap, int *count);
/* Engine statistics since creation (no reference counterpart): graph
 * replays, their GPU time in ms (HIP events around each replay) and the inner
 * row-iterations they advanced (sum over systems of n x inner CG iterations,
 * as seen in the state copies). */
int hfh_batch_stats(const hfh_batch *B, long long *replays, double *busy_ms, long long *row_iters);
/* Copy the solution (x nullable) and outcome of a finished system and free
 * its slot and memory.  Returns HFH_ERR_NOT_CONVERGED for a failed solve. */
int hfh_batch_collect(hfh_batch *B, int id, double *x, hfh_batch_result *res);
int hfh_batch_usage(const hfh_batch *B, int *active, long long *used, long long *cap);

#ifdef __cplusplus
}
#endif
#endif
