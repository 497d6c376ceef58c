/*
 * hfo.c -- CPU oracle for the HiFlow3 sparse linear-algebra hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP platform and the timed
 * CPU baseline.  Never linked into the product library.  Build with
 * oracle/Makefile (gcc, -ffp-contract=off so that no FMA is formed, matching
 * the reference CPU build which has no -march flag).
 */
#include "hfo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Threads of the row-parallel loops (SpMV, the element-wise updates, the
 * rows of one colour in the multicolour sweeps): rows are independent there,
 * so any thread count gives the same bits.  Dot products, level-ordered
 * sweeps and everything else stay sequential, in the reference's order. */
static int hfo_nt = 1;

void hfo_set_threads(int t)
{
    hfo_nt = t < 1 ? 1 : t;
}

int hfo_get_threads(void)
{
    return hfo_nt;
}

#define T double
#define S d
#define ABS fabs
#include "hfo_t.h"
#undef T
#undef S
#undef ABS

#define T float
#define S f
#define ABS fabsf
#include "hfo_t.h"
#undef T
#undef S
#undef ABS

/* CPU_CSR_lMatrix::TransformFromCOO (lmatrix_csr_cpu.cc:951-1044) */
void hfo_csr_from_coo_d(int nrow, int nnz, const int *rows, const int *cols,
                        const double *data, int *row_ptr, int *col, double *val)
{
    memset(row_ptr, 0, sizeof(int) * (nrow + 1));
    memset(col, 0, sizeof(int) * nnz);
    if (data)
        memset(val, 0, sizeof(double) * nnz);
    for (int i = 0; i != nnz; ++i)
        ++row_ptr[rows[i]];
    row_ptr[nrow] = nnz;
    int *index_row = (int *)malloc(sizeof(int) * (nrow > 0 ? nrow : 1));
    int acum = 0;
    for (int i = 0; i != nrow; ++i) {
        const int cur = row_ptr[i];
        row_ptr[i] = acum;
        index_row[i] = acum;
        acum += cur;
    }
    for (int i = 0; i != nnz; ++i) {
        const int t = index_row[rows[i]];
        col[t] = cols[i];
        if (data)
            val[t] = data[i];
        ++index_row[rows[i]];
    }
    free(index_row);
    /* selection sort of the columns of each row (values move with them) */
    for (int k = 0; k != nrow; ++k) {
        const int ie = row_ptr[k + 1];
        for (int i = row_ptr[k]; i != ie; ++i) {
            int mi = i, mc = col[i];
            for (int j = i + 1; j < ie; ++j)
                if (col[j] < mc) {
                    mi = j;
                    mc = col[j];
                }
            int ti = col[i];
            col[i] = col[mi];
            col[mi] = ti;
            if (val) {
                double tv = val[i];
                val[i] = val[mi];
                val[mi] = tv;
            }
        }
    }
}

/* CPU_CSR_lMatrix::add_values (lmatrix_csr_cpu.cc:1147-1180): merge of the
 * sorted column list `cols` against each row's sorted pattern. */
void hfo_csr_add_values_d(int nrow, const int *row_ptr, const int *col, double *val,
                          const int *rows, int nr, const int *cols, int nc,
                          const double *values)
{
    (void)nrow;
    for (int i = 0; i != nr; ++i) {
        int k = 0;
        const int off = i * nc;
        for (int j = row_ptr[rows[i]], je = row_ptr[rows[i] + 1]; k < nc && j != je; ++j) {
            if (cols[k] == col[j]) {
                val[j] += values[off + k];
                ++k;
            } else if (cols[k] < col[j]) {
                while (k < nc && cols[k] < col[j])
                    ++k;
                if (k >= nc)
                    break;
                if (cols[k] == col[j]) {
                    val[j] += values[off + k];
                    ++k;
                }
            }
        }
    }
}

/* CPU_CSR_lMatrix::Reorder (lmatrix_csr_cpu.cc:1847-1862): CSR -> COO,
 * relabel row and column o -> ind[o] with ind[perm[i]] = i
 * (lmatrix_coo_cpu.cc:1019-1054), COO -> CSR (sorted columns). */
void hfo_csr_reorder_d(int n, const int *row, const int *col, const double *val,
                       const int *perm, int *prow, int *pcol, double *pval)
{
    const int nnz = row[n];
    int *ind = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    int *cr = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
    int *cc = (int *)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
    for (int i = 0; i < n; ++i)
        ind[perm[i]] = i;
    for (int i = 0; i < n; ++i)
        for (int j = row[i]; j < row[i + 1]; ++j) {
            cr[j] = ind[i];
            cc[j] = ind[col[j]];
        }
    hfo_csr_from_coo_d(n, nnz, cr, cc, val, prow, pcol, pval);
    free(ind);
    free(cr);
    free(cc);
}

/* CPU_lVector::Reorder (lvector_cpu.cc:75-93) */
void hfo_vec_reorder_d(int n, const double *in, const int *perm, double *out)
{
    for (int i = 0; i < n; ++i)
        out[i] = in[perm[i]];
}

/* CPU_CSR_lMatrix::Multicoloring (lmatrix_csr_cpu.cc:2175-2291): node j gets
 * colour 1, then the row's entries are scanned (row length) times and the
 * colour is incremented whenever an off-diagonal neighbour holds the same
 * colour.  The permutation lists the nodes colour by colour, in increasing
 * node order within a colour. */
int hfo_multicoloring(int n, const int *row, const int *col, int *ncolors, int *nc,
                      int *color_sizes, int *perm)
{
    int hicol = 0;
    for (int i = 0; i < n; ++i)
        nc[i] = 0;
    for (int j = 0; j < n; ++j) {
        nc[j]++;
        const int m = row[j], m1 = row[j + 1];
        for (int k = m; k < m1; ++k)
            for (int i = m; i < m1; ++i) {
                const int dep = col[i];
                if (dep != j) {
                    if (nc[dep] == nc[j])
                        ++nc[j];
                    if (nc[j] > hicol)
                        hicol = nc[j];
                }
            }
    }
    *ncolors = hicol;
    for (int k = 0; k < hicol; ++k)
        color_sizes[k] = 0;
    int zero = 0;
    for (int i = 0; i < n; ++i) {
        if (nc[i] == 0 || nc[i] > hicol) {
            ++zero;
            continue;
        }
        ++color_sizes[nc[i] - 1];
    }
    int *colstart = (int *)calloc(hicol + 1, sizeof(int));
    for (int k = 0; k < hicol; ++k)
        colstart[k + 1] = colstart[k] + color_sizes[k];
    for (int j = 0; j < n; ++j) {
        if (nc[j] == 0 || nc[j] > hicol)
            continue;
        perm[colstart[nc[j] - 1]++] = j;
    }
    free(colstart);
    return zero; /* rows the reference would report as inconsistent */
}

/* CPU_CSR_lMatrix::Levelscheduling (lmatrix_csr_cpu.cc:2295-2391).  Sweep L
 * collects, in increasing row order, the unvisited rows of depth L whose
 * lower-triangular neighbours are all visited (including rows visited earlier
 * in the same sweep), marks them and pushes the depth of their upper
 * neighbours to at least L+1.  The reference only records the level count
 * when a sweep starts with every row done; if the last row is placed in the
 * last possible sweep (L = n-1) it reports 0 levels.  Both values are
 * returned. */
int hfo_levelscheduling(int n, const int *row, const int *col, int *nlevels_reported,
                        int *level_sizes, int *perm)
{
    int *depth = (int *)calloc(n > 0 ? n : 1, sizeof(int));
    char *seen = (char *)calloc(n > 0 ? n : 1, 1);
    int *level_end = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    int done = 0, nlev_rep = 0, nlev = 0;
    for (int L = 0; L < n; ++L) {
        if (done == n) {
            nlev_rep = nlev;
            break;
        }
        for (int k = 0; k < n; ++k) {
            if (seen[k] || depth[k] != L)
                continue;
            int stop = 0;
            for (int j = row[k]; j < row[k + 1]; ++j)
                if (col[j] < k && !seen[col[j]]) {
                    stop = 1;
                    break;
                }
            if (stop)
                continue;
            perm[done++] = k;
            seen[k] = 1;
            for (int j = row[k]; j < row[k + 1]; ++j)
                if (col[j] > k && depth[col[j]] <= depth[k])
                    depth[col[j]] = depth[k] + 1;
        }
        level_end[nlev++] = done;
    }
    int sum = 0;
    for (int i = 0; i < nlev; ++i) {
        level_sizes[i] = level_end[i] - sum;
        sum = level_end[i];
    }
    /* sweeps after the last row add empty levels only when the loop ran on;
     * the computed level count is the number of sweeps that placed rows */
    int nonempty = nlev;
    while (nonempty > 0 && level_sizes[nonempty - 1] == 0)
        --nonempty;
    *nlevels_reported = nlev_rep;
    free(depth);
    free(seen);
    free(level_end);
    return nonempty;
}

/* lPreconditioner_MultiColoring::dropoffs (lpreconditioner_mc.cc:610-662) */
int hfo_mc_dropoffs_d(int n, int *row, int *col, double *val, int ncolors,
                      const int *color_sizes)
{
    (void)row;
    int ndel = 0;
    for (int i = 0; i < n; ++i)
        for (int j = row[i]; j < row[i + 1]; ++j) {
            int off = 0;
            for (int ic = 0; ic < ncolors; ++ic) {
                if (ic >= 1)
                    off += color_sizes[ic - 1];
                if (i >= off && i < off + color_sizes[ic] && col[j] >= off &&
                    col[j] < off + color_sizes[ic] && i != col[j]) {
                    val[j] = 0.0;
                    ++ndel;
                }
            }
        }
    return ndel;
}

/* CoupledVector::CastInteriorFrom (coupled_vector.cc:755-778) */
void hfo_cast_d2f(int n, const double *in, float *out)
{
    for (int i = 0; i < n; ++i)
        out[i] = (float)in[i];
}
void hfo_cast_f2d(int n, const float *in, double *out)
{
    for (int i = 0; i < n; ++i)
        out[i] = (double)in[i];
}

/* MPIR<LAD_high, LAD_low>::Solve and ContinueIteration (mpir.cc:69-327).
 * res0 is reset to the residual just computed after every correction
 * (mpir.cc:275), so the relative tests (MPIR's own and the error-correction
 * gates) compare the residual with itself: with rel_tol < 1 they never end the
 * loop, and only abs_tol, div_tol or the loop limits do.  The low precision
 * tolerances are members of type DataType_low (mpir.h:225-240), so they are
 * rounded to float before the gates compare them and before they reach the
 * inner solver's InitControl (mpir.cc:456-472). */
int hfo_mpir(int n, const int *row, const int *col, const double *val_d,
             const float *val_f, const double *b, double *x,
             const hfo_prec_f *Pl, const hfo_prec_d *Ph,
             const hfo_mpir_params *p, hfo_mpir_result *out, double *hist)
{
    double *resh = (double *)calloc(n, sizeof(double));
    double *corh = (double *)calloc(n, sizeof(double));
    float *resl = (float *)calloc(n, sizeof(float));
    float *corl = (float *)calloc(n, sizeof(float));
    int n_ecl = 0, n_ech = 0, iters = 0, nloop = 0;
    double res, res0;
    hfo_ctrl ecl = p->ecl;
    ecl.atol = (double)(float)ecl.atol;
    ecl.rtol = (double)(float)ecl.rtol;
    ecl.dtol = (double)(float)ecl.dtol;

    hfo_spmv_d(n, row, col, val_d, x, resh);
    hfo_scaleadd_d(n, resh, -1.0, b);
    res = hfo_nrm2_d(n, resh);
    res0 = res;
    if (hist)
        hist[nloop] = res;
#define ECL_OK (n_ecl < p->max_ecl && res > ecl.atol && res > res0 * ecl.rtol && \
                res < ecl.dtol)
#define ECH_OK (n_ech < p->max_ech && res > p->ech.atol && res > res0 * p->ech.rtol && \
                res < p->ech.dtol)
    while (res > p->abs_tol && res > res0 * p->rel_tol && res < p->div_tol &&
           (ECH_OK || ECL_OK)) {
        if (ECL_OK) {
            hfo_result r;
            ++n_ecl;
            hfo_cast_d2f(n, resh, resl);
            memset(corl, 0, sizeof(float) * n);
            hfo_cg_f(n, row, col, val_f, resl, corl, Pl, &ecl, &r, NULL);
            iters += r.iters;
            hfo_cast_f2d(n, corl, corh);
        } else if (ECH_OK) {
            hfo_result r;
            ++n_ech;
            memset(corh, 0, sizeof(double) * n);
            hfo_cg_d(n, row, col, val_d, resh, corh, Ph, &p->ech, &r, NULL);
            iters += r.iters;
        }
        hfo_axpy_d(n, x, corh, 1.0);
        hfo_spmv_d(n, row, col, val_d, x, resh);
        hfo_scaleadd_d(n, resh, -1.0, b);
        res = hfo_nrm2_d(n, resh);
        res0 = res;
        ++nloop;
        if (hist)
            hist[nloop] = res;
    }
#undef ECL_OK
#undef ECH_OK
    out->ret = (res < p->abs_tol || res < res0 * p->rel_tol) ? HFO_SOLVER_SUCCESS
                                                             : HFO_SOLVER_EXCEEDED;
    out->iters = iters;
    out->n_ecl = n_ecl;
    out->n_ech = n_ech;
    out->res = res;
    free(resh);
    free(corh);
    free(resl);
    free(corl);
    return out->ret;
}
