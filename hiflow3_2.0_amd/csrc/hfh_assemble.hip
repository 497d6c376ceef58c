// hfh_assemble.hip -- the seeded benchmark systems assembled on the device.
//
// The host generators (hfl_problem_fill) build every CSR row from its grid
// position alone: the 2D / 3D Q1 FE Laplacians, the cell-centred FV diffusion
// operator (harmonic-mean faces, 2 k on Dirichlet faces) and the first-order
// upwind convection-diffusion operator.  The same rows are assembled here,
// one thread per row, with the same operations in the same order and every
// product and sum rounded separately (the host build has no FMA), so the
// matrix is bit-identical to the host generator's.  The right-hand side is
// the same counter-based hash (splitmix64, integer arithmetic and an exact
// scaling), so it is bit-identical too.  The FV cell coefficients
// k = exp(sigma g) need exp / log / cos and come from the caller
// (hfl_fv_cells), not from the device's math library.
//
// Rows: counted per row, prefix-summed (hipCUB), then filled.  A call moves
// about 12 nnz + 16 n bytes once.
#include "hfh_internal.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

namespace hfh {
namespace {

enum { kQ1_2D = 1, kQ1_3D = 2, kFV = 3, kUpwind = 4 };

struct GridDev {
    int nx, ny, nz;
};

__device__ __forceinline__ void ijk(const GridDev &g, long long r, int *i, int *j, int *k) {
    *i = static_cast<int>(r % g.nx);
    *j = static_cast<int>((r / g.nx) % g.ny);
    *k = static_cast<int>(r / (static_cast<long long>(g.nx) * g.ny));
}

__device__ __forceinline__ int span(int a, int m) { return 1 + (a > 0) + (a < m - 1); }

__global__ __launch_bounds__(kBlock) void k_row_count(int kind, GridDev g, int n, int *__restrict__ cnt) {
    for (int r = blockIdx.x * kBlock + threadIdx.x; r < n; r += gridDim.x * kBlock) {
        int i, j, k;
        ijk(g, r, &i, &j, &k);
        int c;
        if (kind == kQ1_2D || kind == kQ1_3D) {
            c = span(i, g.nx) * span(j, g.ny) * (g.nz > 1 ? span(k, g.nz) : 1);
        } else {
            c = 1 + (i > 0) + (i < g.nx - 1) + (j > 0) + (j < g.ny - 1);
            if (g.nz > 1) c += (k > 0) + (k < g.nz - 1);
        }
        cnt[r] = c;
    }
}

__device__ __forceinline__ double harmonic(double a, double b) {
    return mul_rn(mul_rn(2.0, a), b) / add_rn(a, b);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill_rows(int kind, GridDev g, int n, double coef, double hq1,
                                                      double h, double eps, double bscale,
                                                      const double *__restrict__ kc, const int *__restrict__ row,
                                                      int *__restrict__ col, T *__restrict__ val) {
    for (int r = blockIdx.x * kBlock + threadIdx.x; r < n; r += gridDim.x * kBlock) {
        int i, j, k;
        ijk(g, r, &i, &j, &k);
        int e = row[r];
        const long long plane = static_cast<long long>(g.nx) * g.ny;
        if (kind == kQ1_2D || kind == kQ1_3D) {
            const bool three_d = kind == kQ1_3D;
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
                        if (three_d)
                            v = dist == 0 ? mul_rn(8.0, hq1) / 3.0
                                          : dist == 1 ? 0.0 : dist == 2 ? -hq1 / 6.0 : -hq1 / 12.0;
                        else
                            v = dist == 0 ? 8.0 / 3.0 : -1.0 / 3.0;
                        col[e] = static_cast<int>(ii + g.nx * (jj + static_cast<long long>(g.ny) * kk));
                        val[e] = static_cast<T>(mul_rn(coef, v));
                        ++e;
                    }
                }
            }
            continue;
        }
        // neighbours in ascending column order: k-1, j-1, i-1 | self | i+1, j+1, k+1
        const long long self = r;
        const long long nb[6] = {self - plane, self - g.nx, self - 1, self + 1, self + g.nx, self + plane};
        const bool in3 = g.nz > 1;
        const bool has[6] = {in3 && k > 0, j > 0, i > 0, i < g.nx - 1, j < g.ny - 1, in3 && k < g.nz - 1};
        const bool axis[6] = {in3, true, true, true, true, in3};
        double off[6] = {0, 0, 0, 0, 0, 0};
        double diag = 0;
        if (kind == kFV) {
            const double ks = kc[r];
            for (int q = 0; q < 6; ++q) {
                if (!axis[q]) continue;
                if (has[q]) {
                    const double t = harmonic(ks, kc[nb[q]]);
                    off[q] = -t;
                    diag = add_rn(diag, t);
                } else {
                    diag = add_rn(diag, mul_rn(2.0, ks));   // Dirichlet ghost at half a cell
                }
            }
        } else {
            const double x = mul_rn(add_rn(static_cast<double>(i), 0.5), h);
            const double y = mul_rn(add_rn(static_cast<double>(j), 0.5), h);
            const double b[3] = {mul_rn(-bscale, add_rn(y, -0.5)), mul_rn(bscale, add_rn(x, -0.5)),
                                 mul_rn(bscale, 0.3)};
            for (int q = 0; q < 6; ++q) {
                if (!axis[q]) continue;
                if (has[q]) {
                    off[q] = -eps;
                    diag = add_rn(diag, eps);
                } else {
                    diag = add_rn(diag, mul_rn(2.0, eps));
                }
            }
            for (int d = 0; d < 3; ++d) {
                const double vh = mul_rn(b[d], h);
                if (d == 2 && !in3) continue;
                if (vh > 0) {
                    diag = add_rn(diag, vh);
                    if (has[2 - d]) off[2 - d] = add_rn(off[2 - d], -vh);
                } else if (vh < 0) {
                    diag = add_rn(diag, -vh);
                    if (has[3 + d]) off[3 + d] = add_rn(off[3 + d], vh);
                }
            }
        }
        for (int q = 0; q < 3; ++q)
            if (has[q]) {
                col[e] = static_cast<int>(nb[q]);
                val[e] = static_cast<T>(off[q]);
                ++e;
            }
        col[e] = static_cast<int>(self);
        val[e] = static_cast<T>(diag);
        ++e;
        for (int q = 3; q < 6; ++q)
            if (has[q]) {
                col[e] = static_cast<int>(nb[q]);
                val[e] = static_cast<T>(off[q]);
                ++e;
            }
    }
}

// splitmix64 finaliser and the counter-based uniform in [-1, 1) of the host
// generator: key = mix(mix(seed ^ stream * C) + index), u = (key >> 11) 2^-53.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_rhs(int n, unsigned long long seed, T *__restrict__ out) {
    const unsigned long long base = mix64(seed ^ (1ull * 0xd1342543de82ef95ull));   // stream 1: right-hand side
    for (int r = blockIdx.x * kBlock + threadIdx.x; r < n; r += gridDim.x * kBlock) {
        const double u = mul_rn(static_cast<double>(mix64(base + static_cast<unsigned long long>(r)) >> 11),
                                0x1.0p-53);
        out[r] = static_cast<T>(add_rn(mul_rn(2.0, u), -1.0));
    }
}

int grid_of(const hfh_ctx *c, int n) {
    const long long nb = (static_cast<long long>(n) + kBlock - 1) / kBlock;
    const long long cap = static_cast<long long>(std::max(c->num_cu, 1)) * 16;
    return static_cast<int>(std::max<long long>(1, std::min(nb, cap)));
}

}  // namespace
}  // namespace hfh

using namespace hfh;

extern "C" int hfh_problem_assemble(hfh_csr *A, int kind, int nx, int ny, int nz, double p0, double p1,
                                    unsigned long long seed, const double *kcell, hfh_vec *rhs) {
    HFH_CHECK(A, HFH_ERR_INVALID, "hfh_problem_assemble: null matrix");
    HFH_CHECK(kind >= kQ1_2D && kind <= kUpwind, HFH_ERR_INVALID, "hfh_problem_assemble: kind %d not in 1..4", kind);
    HFH_CHECK(nx >= 1 && ny >= 1 && nz >= 1, HFH_ERR_INVALID, "hfh_problem_assemble: bad grid %d x %d x %d", nx, ny,
              nz);
    const long long N = static_cast<long long>(nx) * ny * nz;
    HFH_CHECK(N == A->nrow && A->ncol == A->nrow, HFH_ERR_INVALID,
              "hfh_problem_assemble: the grid has %lld rows, the matrix %d x %d", N, A->nrow, A->ncol);
    HFH_CHECK(kind != kFV || kcell, HFH_ERR_INVALID, "hfh_problem_assemble: the FV problem needs the cell coefficients");
    HFH_CHECK(!rhs || (rhs->n == A->nrow && rhs->dtype == A->dtype), HFH_ERR_INVALID,
              "hfh_problem_assemble: rhs size/dtype mismatch");
    hfh_ctx *c = A->ctx;
    hipStream_t s = c->stream;
    const int n = A->nrow;
    const GridDev g{nx, ny, nz};
    // scratch: row counts (n + 1 for the scan), scan storage, coefficients
    std::vector<void *> scratch;
    auto release = [&]() {
        (void)hipStreamSynchronize(s);
        for (void *q : scratch) (void)hipFree(q);
    };
    auto get = [&](void **p, size_t bytes) {
        int rc = dev_alloc(p, bytes);
        if (!rc) scratch.push_back(*p);
        return rc;
    };
    int *cnt = nullptr;
    double *kc = nullptr;
    void *tmp = nullptr;
    int rc;
    if ((rc = get(reinterpret_cast<void **>(&cnt), sizeof(int) * (n + 1)))) return release(), rc;
    const int gr = grid_of(c, n);
    if (hipMemsetAsync(cnt + n, 0, sizeof(int), s) != hipSuccess) {
        release();
        return hip_fail(hipErrorUnknown, "hfh_problem_assemble: clearing the count");
    }
    hipLaunchKernelGGL(k_row_count, dim3(gr), dim3(kBlock), 0, s, kind, g, n, cnt);
    if ((rc = launch_check("k_row_count"))) return release(), rc;
    size_t tmp_bytes = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cnt, A->row, n + 1, s);
    if (e == hipSuccess && (rc = get(&tmp, std::max<size_t>(tmp_bytes, 1)))) return release(), rc;
    if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cnt, A->row, n + 1, s);
    if (e != hipSuccess) {
        release();
        return hip_fail(e, "hfh_problem_assemble: row pointer scan");
    }
    // the pattern has to match the matrix's nnz before anything is written
    A->h_row.assign(n + 1, 0);
    e = hipMemcpyAsync(A->h_row.data(), A->row, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        release();
        return hip_fail(e, "hfh_problem_assemble: row pointer download");
    }
    if (A->h_row[n] != A->nnz) {
        release();
        set_error("hfh_problem_assemble: the grid has %d entries, the matrix was created for %d", A->h_row[n], A->nnz);
        return HFH_ERR_INVALID;
    }
    if (kind == kFV) {
        if ((rc = get(reinterpret_cast<void **>(&kc), sizeof(double) * n))) return release(), rc;
        e = hipMemcpyAsync(kc, kcell, sizeof(double) * n, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) {
            release();
            return hip_fail(e, "hfh_problem_assemble: coefficient upload");
        }
    }
    // the host generator's scalars: Q1 mesh width 1/(nx+1); upwind width 1/nx,
    // velocity scale p1 eps / (h sqrt(0.59)) (cell Peclet p1 at the largest |b|)
    const double hq1 = 1.0 / (nx + 1);
    const double h = 1.0 / nx;
    const double eps = p0;
    const double bscale = kind == kUpwind ? p1 * eps / (h * std::sqrt(0.59)) : 0.0;
    if (A->dtype == HFH_F64)
        hipLaunchKernelGGL(k_fill_rows<double>, dim3(gr), dim3(kBlock), 0, s, kind, g, n, p0, hq1, h, eps, bscale,
                           kc, A->row, A->col, static_cast<double *>(A->val));
    else
        hipLaunchKernelGGL(k_fill_rows<float>, dim3(gr), dim3(kBlock), 0, s, kind, g, n, p0, hq1, h, eps, bscale,
                           kc, A->row, A->col, static_cast<float *>(A->val));
    if ((rc = launch_check("k_fill_rows"))) return release(), rc;
    if (rhs) {
        if (A->dtype == HFH_F64)
            hipLaunchKernelGGL(k_rhs<double>, dim3(gr), dim3(kBlock), 0, s, n, seed, static_cast<double *>(rhs->d));
        else
            hipLaunchKernelGGL(k_rhs<float>, dim3(gr), dim3(kBlock), 0, s, n, seed, static_cast<float *>(rhs->d));
        if ((rc = launch_check("k_rhs"))) return release(), rc;
    }
    // host mirror of the columns (the SpMV layout and the setup analyses read it)
    A->h_col.resize(static_cast<size_t>(A->nnz));
    if (A->nnz) {
        e = hipMemcpyAsync(A->h_col.data(), A->col, sizeof(int) * static_cast<size_t>(A->nnz), hipMemcpyDeviceToHost,
                           s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            release();
            return hip_fail(e, "hfh_problem_assemble: column download");
        }
    }
    release();
    A->has_structure = true;
    A->sorted = true;
    if ((rc = spmv_refresh_structure(A))) return rc;
    HFH_HIP(hipStreamSynchronize(s));
    return HFH_OK;
}
