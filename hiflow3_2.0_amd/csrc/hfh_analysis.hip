// hfh_analysis.hip -- preconditioner setup on the device: level schedules
// (CPU_CSR_lMatrix::Levelscheduling, lmatrix_csr_cpu.cc:2295-2391), first-fit
// multicolouring (CPU_CSR_lMatrix::Multicoloring, lmatrix_csr_cpu.cc:2175-2291)
// and the symmetric permutation (CPU_CSR_lMatrix::Reorder,
// lmatrix_csr_cpu.cc:1847-1862 -> lmatrix_coo_cpu.cc:1019-1054).
//
// On the host both analyses are one sequential pass over the rows:
//   level(i)  = 1 + max level(d) over the lower neighbours d < i of row i
//               (0 without one; the upper schedule mirrors it with d > i);
//   colour(i) = the smallest colour >= 1 that no lower neighbour holds.
// Each is the unique fixed point of a map that reads only lower neighbours:
// rows of level 0 are final after one application of the map, rows of level
// k after k + 1.  The device applies the map to every row at once (one
// double-buffered sweep over the CSR arrays) until a sweep changes nothing,
// then orders the rows by (level, index) or (colour, index) with a stable
// radix sort.  Levels, colours and orders are bit-identical to the host
// passes.  Cost: (depth + 1) sweeps of about 4 nnz + 12 n bytes plus the
// neighbour gathers, then a sort of n keys.
#include "hfh_internal.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

namespace hfh {
namespace {

constexpr int kChunk = 16;   // sweeps between convergence checks

// One sweep of the lower map.  flags[0] is set when a value changed (checked
// sweeps only), flags[1] when a row holds an off-diagonal entry (first sweep
// only); one store per workgroup.
template <bool COLOR>
__global__ __launch_bounds__(kBlock) void k_relax_lower(int n, const int *__restrict__ row,
                                                        const int *__restrict__ col, const int *__restrict__ lev_in,
                                                        int *__restrict__ lev_out, const int *__restrict__ clr_in,
                                                        int *__restrict__ clr_out, int *flags, int check, int first) {
    bool changed = false, off = false;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const int j0 = row[i], j1 = row[i + 1];
        int l = 0;
        unsigned long long held = 0;   // bit c-1: colour c held by a lower neighbour (c <= 64)
        for (int q = j0; q < j1; ++q) {
            const int d = col[q];
            off |= d != i;
            if (d < i) {
                l = max(l, lev_in[d] + 1);
                if (COLOR) {
                    const int c = clr_in[d];
                    if (c <= 64) held |= 1ull << (c - 1);
                }
            }
        }
        changed |= l != lev_in[i];
        lev_out[i] = l;
        if (COLOR) {
            int c;
            if (~held != 0ull) {
                c = __builtin_ctzll(~held) + 1;
            } else {   // colours 1..64 all held: the first free one above
                c = 65;
                while (true) {
                    bool taken = false;
                    for (int q = j0; q < j1 && !taken; ++q) taken = col[q] < i && clr_in[col[q]] == c;
                    if (!taken) break;
                    ++c;
                }
            }
            changed |= c != clr_in[i];
            clr_out[i] = c;
        }
    }
    const int any_changed = __syncthreads_or(changed);
    const int any_off = __syncthreads_or(off);
    if (threadIdx.x == 0) {
        if (check && any_changed) flags[0] = 1;
        if (first && any_off) flags[1] = 1;
    }
}

__global__ __launch_bounds__(kBlock) void k_relax_upper(int n, const int *__restrict__ row,
                                                        const int *__restrict__ col, const int *__restrict__ lev_in,
                                                        int *__restrict__ lev_out, int *flags, int check) {
    bool changed = false;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        int l = 0;
        for (int q = row[i]; q < row[i + 1]; ++q) {
            const int d = col[q];
            if (d > i) l = max(l, lev_in[d] + 1);
        }
        changed |= l != lev_in[i];
        lev_out[i] = l;
    }
    const int any_changed = __syncthreads_or(changed);
    if (threadIdx.x == 0 && check && any_changed) flags[0] = 1;
}

__global__ __launch_bounds__(kBlock) void k_fill_int(int n, int *__restrict__ a, int v) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) a[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_iota(int n, int *__restrict__ a) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) a[i] = i;
}

// ptr[k - base] = first position of key k in the sorted keys (every key in
// [base, base + nb) occurs; empty buckets get the next start all the same).
__global__ __launch_bounds__(kBlock) void k_bucket_starts(int n, const int *__restrict__ key, int base,
                                                          int *__restrict__ ptr) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const int k = key[i] - base;
        const int prev = i == 0 ? -1 : key[i - 1] - base;
        for (int b = prev + 1; b <= k; ++b) ptr[b] = i;
    }
}

__global__ __launch_bounds__(kBlock) void k_max_int(int n, const int *__restrict__ a, int *out) {
    __shared__ int sh[kBlock];
    int m = 0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) m = max(m, a[i]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (static_cast<int>(threadIdx.x) < s) sh[threadIdx.x] = max(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, sh[0]);
}

// Row i of the permuted matrix is row perm[i] with its columns renumbered by
// inv and sorted ascending (insertion sort: rows are short).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_permute_rows(int n, const int *__restrict__ row, const int *__restrict__ col,
                                                         const T *__restrict__ val, const int *__restrict__ perm,
                                                         const int *__restrict__ inv, const int *__restrict__ prow,
                                                         int *__restrict__ pcol, T *__restrict__ pval) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const int o = perm[i], s = row[o], len = row[o + 1] - s, b = prow[i];
        for (int q = 0; q < len; ++q) {
            const int c = inv[col[s + q]];
            const T v = val[s + q];
            int p = q;
            while (p > 0 && pcol[b + p - 1] > c) {
                pcol[b + p] = pcol[b + p - 1];
                pval[b + p] = pval[b + p - 1];
                --p;
            }
            pcol[b + p] = c;
            pval[b + p] = v;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_inverse(int n, const int *__restrict__ perm, int *__restrict__ inv) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) inv[perm[i]] = i;
}

// Device buffers released together, after the stream's queued work.
struct Bufs {
    hipStream_t s;
    std::vector<void *> p;
    explicit Bufs(hipStream_t st) : s(st) {}
    ~Bufs() {
        if (!p.empty()) (void)hipStreamSynchronize(s);
        for (void *q : p)
            if (q) (void)hipFree(q);
    }
    int get(void **out, size_t bytes) {
        int rc = dev_alloc(out, bytes);
        if (!rc) p.push_back(*out);
        return rc;
    }
};

int grid_for(const hfh_ctx *c, int n) {
    const long long nb = (static_cast<long long>(n) + kBlock - 1) / kBlock;
    const long long cap = static_cast<long long>(std::max(c->num_cu, 1)) * 16;
    return static_cast<int>(std::max<long long>(1, std::min(nb, cap)));
}

// Iterates the map to its fixed point.  On return lev (and clr) hold the
// result; *has_off tells whether any row has an off-diagonal entry.
int relax(const hfh_csr *A, bool upper, bool color, Bufs &B, int **lev, int **clr, bool *has_off) {
    hfh_ctx *c = A->ctx;
    hipStream_t s = c->stream;
    const int n = A->nrow;
    int *L[2] = {nullptr, nullptr}, *K[2] = {nullptr, nullptr}, *flags = nullptr;
    int rc;
    if ((rc = B.get(reinterpret_cast<void **>(&L[0]), sizeof(int) * n)) ||
        (rc = B.get(reinterpret_cast<void **>(&L[1]), sizeof(int) * n)) ||
        (rc = B.get(reinterpret_cast<void **>(&flags), sizeof(int) * 2)))
        return rc;
    if (color && ((rc = B.get(reinterpret_cast<void **>(&K[0]), sizeof(int) * n)) ||
                  (rc = B.get(reinterpret_cast<void **>(&K[1]), sizeof(int) * n))))
        return rc;
    const int g = grid_for(c, n);
    HFH_HIP(hipMemsetAsync(L[0], 0, sizeof(int) * n, s));
    HFH_HIP(hipMemsetAsync(flags, 0, sizeof(int) * 2, s));
    if (color) hipLaunchKernelGGL(k_fill_int, dim3(g), dim3(kBlock), 0, s, n, K[0], 1);
    if ((rc = launch_check("k_fill_int"))) return rc;
    int cur = 0;
    int *h = reinterpret_cast<int *>(c->host_res);   // pinned, at least 2 ints
    // the fixed point is reached after depth + 1 <= n sweeps
    for (long long done = 0; done <= n + kChunk; done += kChunk) {
        for (int k = 0; k < kChunk; ++k) {
            const int check = k == kChunk - 1, first = done == 0 && k == 0;
            if (check) HFH_HIP(hipMemsetAsync(flags, 0, sizeof(int), s));
            if (upper)
                hipLaunchKernelGGL(k_relax_upper, dim3(g), dim3(kBlock), 0, s, n, A->row, A->col, L[cur], L[cur ^ 1],
                                   flags, check);
            else if (color)
                hipLaunchKernelGGL(k_relax_lower<true>, dim3(g), dim3(kBlock), 0, s, n, A->row, A->col, L[cur],
                                   L[cur ^ 1], K[cur], K[cur ^ 1], flags, check, first);
            else
                hipLaunchKernelGGL(k_relax_lower<false>, dim3(g), dim3(kBlock), 0, s, n, A->row, A->col, L[cur],
                                   L[cur ^ 1], nullptr, nullptr, flags, check, first);
            if ((rc = launch_check("k_relax"))) return rc;
            cur ^= 1;
        }
        HFH_HIP(hipMemcpyAsync(h, flags, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
        HFH_HIP(hipStreamSynchronize(s));
        if (h[0] == 0) {
            *lev = L[cur];
            if (clr) *clr = color ? K[cur] : nullptr;
            if (has_off) *has_off = h[1] != 0;
            return HFH_OK;
        }
    }
    set_error("device analysis: no fixed point after %d sweeps", n + 2 * kChunk);
    return HFH_ERR_INVALID;
}

// Stable order of the rows by key (key in [base, base + nb)): rows_out (device)
// and the bucket pointer on the host (nb + 1 entries).
int bucket(hfh_ctx *c, int n, const int *key, int base, Bufs &B, int **rows_out, std::vector<int> &ptr) {
    hipStream_t s = c->stream;
    const int g = grid_for(c, n);
    int *mx = nullptr, *vin = nullptr, *vout = nullptr, *kout = nullptr, *dptr = nullptr;
    int rc;
    if ((rc = B.get(reinterpret_cast<void **>(&mx), sizeof(int))) ||
        (rc = B.get(reinterpret_cast<void **>(&vin), sizeof(int) * n)) ||
        (rc = B.get(reinterpret_cast<void **>(&vout), sizeof(int) * n)) ||
        (rc = B.get(reinterpret_cast<void **>(&kout), sizeof(int) * n)))
        return rc;
    HFH_HIP(hipMemsetAsync(mx, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_max_int, dim3(std::min(g, 1024)), dim3(kBlock), 0, s, n, key, mx);
    hipLaunchKernelGGL(k_iota, dim3(g), dim3(kBlock), 0, s, n, vin);
    if ((rc = launch_check("k_iota"))) return rc;
    int *h = reinterpret_cast<int *>(c->host_res);
    HFH_HIP(hipMemcpyAsync(h, mx, sizeof(int), hipMemcpyDeviceToHost, s));
    HFH_HIP(hipStreamSynchronize(s));
    const int kmax = h[0];
    const int nb = kmax - base + 1;
    int bits = 1;
    while (bits < 31 && (1LL << bits) <= kmax) ++bits;
    size_t tmp_bytes = 0;
    HFH_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key, kout, vin, vout, n, 0, bits, s));
    void *tmp = nullptr;
    if ((rc = B.get(&tmp, std::max<size_t>(tmp_bytes, 1)))) return rc;
    HFH_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, key, kout, vin, vout, n, 0, bits, s));
    if ((rc = B.get(reinterpret_cast<void **>(&dptr), sizeof(int) * (nb + 1)))) return rc;
    hipLaunchKernelGGL(k_bucket_starts, dim3(g), dim3(kBlock), 0, s, n, kout, base, dptr);
    if ((rc = launch_check("k_bucket_starts"))) return rc;
    ptr.assign(nb + 1, n);
    HFH_HIP(hipMemcpyAsync(ptr.data(), dptr, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
    HFH_HIP(hipStreamSynchronize(s));
    ptr[nb] = n;
    *rows_out = vout;
    return HFH_OK;
}

int check_square(const hfh_csr *A, const char *what) {
    HFH_CHECK(A && A->has_structure, HFH_ERR_INVALID, "%s: matrix has no structure", what);
    HFH_CHECK(A->nrow == A->ncol && A->nrow > 0, HFH_ERR_INVALID, "%s: needs a non-empty square matrix", what);
    return HFH_OK;
}

}  // namespace
}  // namespace hfh

using namespace hfh;

extern "C" int hfh_csr_levels(const hfh_csr *A, int upper, int *nlev, int *level_ptr, int *level_rows) {
    int rc = check_square(A, "hfh_csr_levels");
    if (rc) return rc;
    HFH_CHECK(nlev && level_ptr && level_rows, HFH_ERR_INVALID, "hfh_csr_levels: null output");
    Bufs B(A->ctx->stream);
    int *lev = nullptr, *rows = nullptr;
    if ((rc = relax(A, upper != 0, false, B, &lev, nullptr, nullptr))) return rc;
    std::vector<int> ptr;
    if ((rc = bucket(A->ctx, A->nrow, lev, 0, B, &rows, ptr))) return rc;
    HFH_HIP(hipMemcpyAsync(level_rows, rows, sizeof(int) * A->nrow, hipMemcpyDeviceToHost, A->ctx->stream));
    HFH_HIP(hipStreamSynchronize(A->ctx->stream));
    *nlev = static_cast<int>(ptr.size()) - 1;
    std::copy(ptr.begin(), ptr.end(), level_ptr);
    return HFH_OK;
}

extern "C" int hfh_csr_multicoloring(const hfh_csr *A, int *ncolors, int *color_of, int *color_sizes, int *perm) {
    int rc = check_square(A, "hfh_csr_multicoloring");
    if (rc) return rc;
    HFH_CHECK(ncolors && color_sizes && perm, HFH_ERR_INVALID, "hfh_csr_multicoloring: null output");
    Bufs B(A->ctx->stream);
    int *lev = nullptr, *clr = nullptr, *rows = nullptr;
    bool off = false;
    if ((rc = relax(A, false, true, B, &lev, &clr, &off))) return rc;
    // the reference counts only colours met at rows with off-diagonal
    // entries; a pattern without one has no colouring
    HFH_CHECK(off, HFH_ERR_INVALID, "hfh_csr_multicoloring: the pattern has no off-diagonal entry");
    std::vector<int> ptr;
    if ((rc = bucket(A->ctx, A->nrow, clr, 1, B, &rows, ptr))) return rc;
    hipStream_t s = A->ctx->stream;
    HFH_HIP(hipMemcpyAsync(perm, rows, sizeof(int) * A->nrow, hipMemcpyDeviceToHost, s));
    if (color_of) HFH_HIP(hipMemcpyAsync(color_of, clr, sizeof(int) * A->nrow, hipMemcpyDeviceToHost, s));
    HFH_HIP(hipStreamSynchronize(s));
    *ncolors = static_cast<int>(ptr.size()) - 1;
    for (int k = 0; k < *ncolors; ++k) color_sizes[k] = ptr[k + 1] - ptr[k];
    return HFH_OK;
}

extern "C" int hfh_csr_permute(const hfh_csr *A, const int *perm, hfh_csr *P) {
    int rc = check_square(A, "hfh_csr_permute");
    if (rc) return rc;
    HFH_CHECK(perm && P, HFH_ERR_INVALID, "hfh_csr_permute: null argument");
    HFH_CHECK(P != A && P->nrow == A->nrow && P->ncol == A->ncol && P->nnz == A->nnz && P->dtype == A->dtype,
              HFH_ERR_INVALID, "hfh_csr_permute: output must be another matrix of the same dims and dtype");
    const int n = A->nrow;
    // perm[new] = old must be a permutation (it indexes device arrays)
    std::vector<int> prow(n + 1, 0), seen(n, 0);
    for (int i = 0; i < n; ++i) {
        HFH_CHECK(perm[i] >= 0 && perm[i] < n && !seen[perm[i]], HFH_ERR_INVALID,
                  "hfh_csr_permute: perm is not a permutation (entry %d)", i);
        seen[perm[i]] = 1;
        prow[i + 1] = prow[i] + (A->h_row[perm[i] + 1] - A->h_row[perm[i]]);
    }
    hfh_ctx *c = A->ctx;
    hipStream_t s = c->stream;
    Bufs B(A->ctx->stream);
    int *dperm = nullptr, *dinv = nullptr;
    if ((rc = B.get(reinterpret_cast<void **>(&dperm), sizeof(int) * n)) ||
        (rc = B.get(reinterpret_cast<void **>(&dinv), sizeof(int) * n)))
        return rc;
    HFH_HIP(hipMemcpyAsync(dperm, perm, sizeof(int) * n, hipMemcpyHostToDevice, s));
    HFH_HIP(hipMemcpyAsync(P->row, prow.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    const int g = grid_for(c, n);
    hipLaunchKernelGGL(k_inverse, dim3(g), dim3(kBlock), 0, s, n, dperm, dinv);
    if (A->dtype == HFH_F64)
        hipLaunchKernelGGL(k_permute_rows<double>, dim3(g), dim3(kBlock), 0, s, n, A->row, A->col,
                           (const double *)A->val, dperm, dinv, P->row, P->col, (double *)P->val);
    else
        hipLaunchKernelGGL(k_permute_rows<float>, dim3(g), dim3(kBlock), 0, s, n, A->row, A->col,
                           (const float *)A->val, dperm, dinv, P->row, P->col, (float *)P->val);
    if ((rc = launch_check("k_permute_rows"))) return rc;
    // host mirrors of the pattern (the SpMV layout and the sweep builders read them)
    P->h_row = prow;
    P->h_col.resize(static_cast<size_t>(A->nnz));
    if (A->nnz)
        HFH_HIP(hipMemcpyAsync(P->h_col.data(), P->col, sizeof(int) * static_cast<size_t>(A->nnz),
                               hipMemcpyDeviceToHost, s));
    HFH_HIP(hipStreamSynchronize(s));
    P->has_structure = true;
    P->sorted = true;
    return spmv_refresh_structure(P);
}
