// hflmp_capi.cc -- C entry points of the host layer (problem generation and
// preconditioner setup) for bindings that cannot use the C++ classes.
#include "../../include/hflmp.h"

#include <cstddef>
#include <cstdio>
#include <vector>

#include "analysis.h"
#include "problems.h"

using namespace hflmp;

static thread_local char g_err[512] = "";

void hflmp_set_error(const char *msg) { snprintf(g_err, sizeof g_err, "%s", msg); }

extern "C" const char *hfl_last_error(void) { return g_err; }

static ProblemSpec spec(int kind, int nx, int ny, int nz, double p0, double p1, unsigned long long seed) {
    return ProblemSpec{kind, nx, ny, nz, p0, p1, static_cast<uint64_t>(seed)};
}

extern "C" int hfl_problem_size(int kind, int nx, int ny, int nz, double p0, double p1, unsigned long long seed,
                                int *n, long long *nnz) {
    if (!n || !nnz) return -1;
    return problem_size(spec(kind, nx, ny, nz, p0, p1, seed), n, nnz);
}

extern "C" int hfl_problem_fill(int kind, int nx, int ny, int nz, double p0, double p1, unsigned long long seed,
                                int *row, int *col, double *val, double *rhs) {
    if (!row || !col || !val) return -1;
    return problem_fill(spec(kind, nx, ny, nz, p0, p1, seed), row, col, val, rhs);
}

extern "C" int hfl_fv_cells(int nx, int ny, int nz, double sigma, unsigned long long seed, double *kcell,
                            double *rhs) {
    return fv_cells(spec(kFvDiffusion, nx, ny, nz, sigma, 0.0, seed), kcell, rhs);
}

extern "C" double hfl_uniform_pm1(unsigned long long seed, unsigned long long stream, unsigned long long index) {
    return uniform_pm1(seed, stream, index);
}

extern "C" double hfl_normal01(unsigned long long seed, unsigned long long stream, unsigned long long index) {
    return normal01(seed, stream, index);
}

extern "C" int hfl_multicoloring(int n, const int *row, const int *col, int *ncolors, int *color_of,
                                 int *color_sizes, int *perm) {
    if (n < 0 || !row || !col || !ncolors || !color_of || !color_sizes || !perm) return -1;
    std::vector<int> sizes;
    const int rc = multicoloring(n, row, col, ncolors, color_of, sizes, perm);
    if (rc) return rc;
    for (int c = 0; c < *ncolors; ++c) color_sizes[c] = sizes[c];
    return 0;
}

static int levels(bool upper, int n, const int *row, const int *col, int *nlev, int *level_ptr, int *level_rows) {
    if (n < 0 || !row || !col || !nlev || !level_ptr || !level_rows) return -1;
    std::vector<int> ptr;
    const int rc = upper ? levels_upper(n, row, col, ptr, level_rows) : levels_lower(n, row, col, ptr, level_rows);
    if (rc) return rc;
    *nlev = static_cast<int>(ptr.size()) - 1;
    for (size_t l = 0; l < ptr.size(); ++l) level_ptr[l] = ptr[l];
    return 0;
}

extern "C" int hfl_levels_lower(int n, const int *row, const int *col, int *nlev, int *level_ptr, int *level_rows) {
    return levels(false, n, row, col, nlev, level_ptr, level_rows);
}

extern "C" int hfl_levels_upper(int n, const int *row, const int *col, int *nlev, int *level_ptr, int *level_rows) {
    return levels(true, n, row, col, nlev, level_ptr, level_rows);
}

extern "C" int hfl_ilu0(int n, const int *row, const int *col, double *val) {
    if (n < 0 || !row || !col || !val) return -1;
    return ilu0(n, row, col, val);
}

extern "C" int hfl_csr_permute(int n, const int *row, const int *col, const double *val, const int *perm, int *prow,
                               int *pcol, double *pval) {
    if (n < 0 || !row || !col || !val || !perm || !prow || !pcol || !pval) return -1;
    return csr_permute(n, row, col, val, perm, prow, pcol, pval);
}
