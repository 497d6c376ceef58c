// analysis.h -- host-side setup of the preconditioners: multicolouring,
// level schedules, ILU(0) factorisation and symmetric CSR permutation.
// The reference runs these on the CPU as well (lpreconditioner_mc.cc:157-300,
// lpreconditioner.cc:388-411); only the applies run on the device.
#pragma once

#include <vector>

namespace hflmp {

// CPU_CSR_lMatrix::Multicoloring (lmatrix_csr_cpu.cc:2175-2291): greedy
// first-fit colouring in natural row order.  color_of[i] is 1-based; perm
// lists rows colour by colour in increasing row order (perm[new] = old).
int multicoloring(int n, const int *row, const int *col, int *ncolors, int *color_of,
                  std::vector<int> &color_sizes, int *perm);

// Level schedule of the forward (lower-triangular) sweep: level(i) = 1 + max
// level of its lower neighbours (0 without any), rows ordered by (level,
// row).  For structurally symmetric patterns these are exactly the level
// sets of CPU_CSR_lMatrix::Levelscheduling (lmatrix_csr_cpu.cc:2295-2391).
int levels_lower(int n, const int *row, const int *col, std::vector<int> &level_ptr, int *level_rows);
// The same for the backward (upper-triangular) sweep, levels counted from
// the last row.
int levels_upper(int n, const int *row, const int *col, std::vector<int> &level_ptr, int *level_rows);

// CPU_CSR_lMatrix::ilu0 (lmatrix_csr_cpu.cc:1880-1930): IKJ ILU(0) in place on
// a column-sorted pattern.  Returns the number of missing diagonal entries.
int ilu0(int n, const int *row, const int *col, double *val);

// CPU_CSR_lMatrix::Reorder (lmatrix_csr_cpu.cc:1847-1862): B = P A P^T with
// perm[new] = old; columns sorted ascending in each row.
int csr_permute(int n, const int *row, const int *col, const double *val, const int *perm, int *prow,
                int *pcol, double *pval);

}  // namespace hflmp

namespace hflmp {

// CPU_CSR_lMatrix::Levelscheduling (lmatrix_csr_cpu.cc:2295-2391) itself:
// sweeps L = 0, 1, ... over the rows in order; a row of depth L whose lower
// neighbours are all placed is placed, and raises the depth of its higher
// neighbours to L + 1.  *nlevels is the count the reference reports: the
// levels are only counted when a sweep starts with every row placed, so a
// schedule whose last row is placed in sweep n - 1 reports 0 levels.
// level_sizes needs n entries (*nlevels are written), perm n.
void levelscheduling(int n, const int *row, const int *col, int *nlevels, int *level_sizes, int *perm);

}  // namespace hflmp
