// problems.h -- seeded synthetic linear systems in CSR form (host side).
//
// The reference assembles these systems with its FE / FV modules; here they
// are generated directly as CSR with int32 indices, rows in lexicographic
// order (x fastest) and columns ascending within a row
// (lmatrix_csr_cpu.cc:1017-1044).  Random inputs come from a counter-based
// generator keyed by (seed, stream, index), so any entry can be recomputed
// independently of the others and of the thread that makes it.
#pragma once

#include <cstdint>

namespace hflmp {

enum ProblemKind {
    kQ1Laplace2D = 1,   // 2D Q1 bilinear FE Laplacian, interior nodes of an N x N cell mesh
    kQ1Laplace3D = 2,   // 3D Q1 trilinear FE Laplacian, interior nodes of an N^3 cell mesh
    kFvDiffusion = 3,   // cell-centred FV -div(k grad u), 5-pt (nz == 1) or 7-pt, k = exp(sigma g)
    kUpwindConvDiff = 4 // first-order upwind -eps lap u + b.grad u on an N^3 cell grid
};

struct ProblemSpec {
    int kind;
    int nx, ny, nz;
    double p0, p1;          // kind-specific: Q1: coefficient; FV: sigma; upwind: eps, cell Peclet
    uint64_t seed;
};

uint64_t mix64(uint64_t z);
double uniform01(uint64_t seed, uint64_t stream, uint64_t index);     // [0, 1)
double uniform_pm1(uint64_t seed, uint64_t stream, uint64_t index);   // [-1, 1)
double normal01(uint64_t seed, uint64_t stream, uint64_t index);      // Box-Muller

int problem_size(const ProblemSpec &p, int *n, long long *nnz);
int problem_fill(const ProblemSpec &p, int *row, int *col, double *val, double *rhs);
// Cell data of the FV diffusion problem: k_c = exp(p0 g_c) and the
// right-hand side, exactly the values problem_fill uses (either may be null).
int fv_cells(const ProblemSpec &p, double *kc, double *rhs);

}  // namespace hflmp
