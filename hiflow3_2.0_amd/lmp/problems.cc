// problems.cc -- seeded synthetic CSR systems for the benchmark and test
// configurations.  Rows are counted, prefix-summed and then filled row by
// row in parallel; every value depends only on its row, so the output is
// identical for any thread count.
#include "problems.h"

#include <cmath>
#include <vector>

namespace hflmp {

uint64_t mix64(uint64_t z) {
    // splitmix64 finaliser
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static inline uint64_t key(uint64_t seed, uint64_t stream, uint64_t index) {
    return mix64(mix64(seed ^ (stream * 0xd1342543de82ef95ull)) + index);
}

double uniform01(uint64_t seed, uint64_t stream, uint64_t index) {
    return static_cast<double>(key(seed, stream, index) >> 11) * 0x1.0p-53;
}

double uniform_pm1(uint64_t seed, uint64_t stream, uint64_t index) {
    return 2.0 * uniform01(seed, stream, index) - 1.0;
}

double normal01(uint64_t seed, uint64_t stream, uint64_t index) {
    const double u1 = (static_cast<double>(key(seed, stream, 2 * index) >> 11) + 1.0) * 0x1.0p-53;  // (0, 1]
    const double u2 = uniform01(seed, stream, 2 * index + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

namespace {

constexpr uint64_t kStreamRhs = 1, kStreamCoef = 2;

struct Grid {
    int nx, ny, nz;
    long long n() const { return static_cast<long long>(nx) * ny * nz; }
    long long idx(int i, int j, int k) const { return i + static_cast<long long>(nx) * (j + static_cast<long long>(ny) * k); }
};

// Q1 stencils: 2D interior node couples to its 8 neighbours (value -1/3) with
// diagonal 8/3; 3D: centre 8h/3, face neighbours 0 (kept in the pattern),
// edge neighbours -h/6, corner neighbours -h/12.
int q1_count(const Grid &g, long long r) {
    const int i = static_cast<int>(r % g.nx);
    const int j = static_cast<int>((r / g.nx) % g.ny);
    const int k = static_cast<int>(r / (static_cast<long long>(g.nx) * g.ny));
    auto span = [](int a, int m) { return 1 + (a > 0) + (a < m - 1); };
    return span(i, g.nx) * span(j, g.ny) * (g.nz > 1 ? span(k, g.nz) : 1);
}

void q1_row(const Grid &g, long long r, bool three_d, double coef, double h, int *col, double *val) {
    const int i = static_cast<int>(r % g.nx);
    const int j = static_cast<int>((r / g.nx) % g.ny);
    const int k = static_cast<int>(r / (static_cast<long long>(g.nx) * g.ny));
    int e = 0;
    const int kd0 = three_d ? -1 : 0, kd1 = three_d ? 1 : 0;
    for (int dk = kd0; dk <= kd1; ++dk) {
        const int kk = k + dk;
        if (kk < 0 || kk >= g.nz) continue;
        for (int dj = -1; dj <= 1; ++dj) {
            const int jj = j + dj;
            if (jj < 0 || jj >= g.ny) continue;
            for (int di = -1; di <= 1; ++di) {
                const int ii = i + di;
                if (ii < 0 || ii >= g.nx) continue;
                const int dist = (di != 0) + (dj != 0) + (dk != 0);
                double v;
                if (three_d) {
                    v = dist == 0 ? 8.0 * h / 3.0 : dist == 1 ? 0.0 : dist == 2 ? -h / 6.0 : -h / 12.0;
                } else {
                    v = dist == 0 ? 8.0 / 3.0 : -1.0 / 3.0;
                }
                col[e] = static_cast<int>(g.idx(ii, jj, kk));
                val[e] = coef * v;
                ++e;
            }
        }
    }
}

// 5-pt / 7-pt rows: neighbours in ascending index order
// (k-1, j-1, i-1, self, i+1, j+1, k+1).
int fv_count(const Grid &g, long long r) {
    const int i = static_cast<int>(r % g.nx);
    const int j = static_cast<int>((r / g.nx) % g.ny);
    const int k = static_cast<int>(r / (static_cast<long long>(g.nx) * g.ny));
    int c = 1 + (i > 0) + (i < g.nx - 1) + (j > 0) + (j < g.ny - 1);
    if (g.nz > 1) c += (k > 0) + (k < g.nz - 1);
    return c;
}

inline double harmonic(double a, double b) { return 2.0 * a * b / (a + b); }

}  // namespace

int problem_size(const ProblemSpec &p, int *n, long long *nnz) {
    Grid g{p.nx, p.ny, p.nz};
    if (p.nx < 1 || p.ny < 1 || p.nz < 1) return -1;
    if (g.n() > 0x7fffffffLL) return -1;
    long long total = 0;
    const long long N = g.n();
    switch (p.kind) {
    case kQ1Laplace2D:
    case kQ1Laplace3D:
        for (long long r = 0; r < N; ++r) total += q1_count(g, r);
        break;
    case kFvDiffusion:
    case kUpwindConvDiff:
        for (long long r = 0; r < N; ++r) total += fv_count(g, r);
        break;
    default:
        return -1;
    }
    *n = static_cast<int>(N);
    *nnz = total;
    return total > 0x7fffffffLL ? -1 : 0;
}

int fv_cells(const ProblemSpec &p, double *kc, double *rhs) {
    const long long n = static_cast<long long>(p.nx) * p.ny * p.nz;
    if (p.nx < 1 || p.ny < 1 || p.nz < 1 || n > 0x7fffffffLL) return -1;
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < n; ++c) {
        if (kc) kc[c] = std::exp(p.p0 * normal01(p.seed, kStreamCoef, c));
        if (rhs) rhs[c] = uniform_pm1(p.seed, kStreamRhs, c);
    }
    return 0;
}

int problem_fill(const ProblemSpec &p, int *row, int *col, double *val, double *rhs) {
    int n;
    long long nnz;
    if (problem_size(p, &n, &nnz)) return -1;
    Grid g{p.nx, p.ny, p.nz};
    const bool q1 = p.kind == kQ1Laplace2D || p.kind == kQ1Laplace3D;
    row[0] = 0;
    for (long long r = 0; r < n; ++r) row[r + 1] = row[r] + (q1 ? q1_count(g, r) : fv_count(g, r));
    std::vector<double> kc;
    if (p.kind == kFvDiffusion) {
        kc.resize(n);
        fv_cells(p, kc.data(), nullptr);
    }
    // upwind velocity scale: cell Peclet |b| h / eps = p1 at the largest |b0| = sqrt(0.59)
    const double h = 1.0 / g.nx;
    const double eps = p.p0;
    const double bscale = p.kind == kUpwindConvDiff ? p.p1 * eps / (h * std::sqrt(0.59)) : 0.0;
#pragma omp parallel for schedule(static)
    for (long long r = 0; r < n; ++r) {
        int *c = col + row[r];
        double *v = val + row[r];
        if (rhs) rhs[r] = uniform_pm1(p.seed, kStreamRhs, r);
        if (q1) {
            q1_row(g, r, p.kind == kQ1Laplace3D, p.p0, 1.0 / (g.nx + 1), c, v);
            continue;
        }
        const int i = static_cast<int>(r % g.nx);
        const int j = static_cast<int>((r / g.nx) % g.ny);
        const int k = static_cast<int>(r / (static_cast<long long>(g.nx) * g.ny));
        // neighbour offsets in ascending column order with their existence
        const long long nb[6] = {g.idx(i, j, k) - static_cast<long long>(g.nx) * g.ny, g.idx(i, j, k) - g.nx,
                                 g.idx(i, j, k) - 1, g.idx(i, j, k) + 1, g.idx(i, j, k) + g.nx,
                                 g.idx(i, j, k) + static_cast<long long>(g.nx) * g.ny};
        const bool in3 = g.nz > 1;
        const bool has[6] = {in3 && k > 0, j > 0, i > 0, i < g.nx - 1, j < g.ny - 1, in3 && k < g.nz - 1};
        const bool axis_exists[6] = {in3, true, true, true, true, in3};
        double off[6] = {0, 0, 0, 0, 0, 0};
        double diag = 0;
        if (p.kind == kFvDiffusion) {
            const double kself = kc[r];
            for (int q = 0; q < 6; ++q) {
                if (!axis_exists[q]) continue;
                if (has[q]) {
                    const double t = harmonic(kself, kc[nb[q]]);
                    off[q] = -t;
                    diag += t;
                } else {
                    diag += 2.0 * kself;  // Dirichlet ghost at half a cell
                }
            }
        } else {
            const double x = (i + 0.5) * h, y = (j + 0.5) * h;
            const double b[3] = {-bscale * (y - 0.5), bscale * (x - 0.5), bscale * 0.3};
            for (int q = 0; q < 6; ++q) {
                if (!axis_exists[q]) continue;
                if (has[q]) {
                    off[q] = -eps;
                    diag += eps;
                } else {
                    diag += 2.0 * eps;
                }
            }
            // axis d: minus neighbour is q = 2 - d, plus neighbour q = 3 + d
            for (int d = 0; d < 3; ++d) {
                const double vh = b[d] * h;
                if (d == 2 && !in3) continue;
                if (vh > 0) {
                    diag += vh;
                    if (has[2 - d]) off[2 - d] -= vh;
                } else if (vh < 0) {
                    diag -= vh;
                    if (has[3 + d]) off[3 + d] += vh;
                }
            }
        }
        int e = 0;
        for (int q = 0; q < 3; ++q)
            if (has[q]) { c[e] = static_cast<int>(nb[q]); v[e] = off[q]; ++e; }
        c[e] = static_cast<int>(r);
        v[e] = diag;
        ++e;
        for (int q = 3; q < 6; ++q)
            if (has[q]) { c[e] = static_cast<int>(nb[q]); v[e] = off[q]; ++e; }
    }
    return 0;
}

}  // namespace hflmp
