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
// Both depend only on lower neighbours, so the device visits the rows in a
// level-synchronous topological order (Kahn): each row counts its lower
// neighbours; the rows without one form level 0; processing level L gives
// each of its rows its colour (its lower neighbours all sit on earlier
// levels) and decrements the count of every row that depends on it, and a row
// whose count reaches zero joins level L + 1.  One launch per level, O(nnz)
// atomics in all.  The level and colour of every row are deterministic (the
// order inside a frontier is not, and is not used): a stable radix sort by
// (level or colour, row) then yields the host's order.  Levels, colours and
// orders are bit-identical to the host passes.
#include "hfh_internal.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

namespace hfh {
namespace {

constexpr int kChunk = 32;   // level launches between termination checks

__device__ __forceinline__ bool depends(int d, int i, int upper) { return upper ? d > i : d < i; }

// indeg[i] = neighbours row i waits for; depcnt[d] = rows that wait for d.
// flags[0]: some row has an off-diagonal entry (one store per workgroup).
__global__ __launch_bounds__(kBlock) void k_dep_count(int n, const int *__restrict__ row, const int *__restrict__ col,
                                                      int upper, int *__restrict__ indeg, int *depcnt, int *flags) {
    bool off = false;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        int c = 0;
        for (int q = row[i]; q < row[i + 1]; ++q) {
            const int d = col[q];
            off |= d != i;
            if (depends(d, i, upper)) {
                ++c;
                atomicAdd(&depcnt[d], 1);
            }
        }
        indeg[i] = c;
    }
    if (__syncthreads_or(off) && threadIdx.x == 0) flags[0] = 1;
}

// deps[depptr[d] ...] = the rows that wait for d (in no particular order).
__global__ __launch_bounds__(kBlock) void k_dep_fill(int n, const int *__restrict__ row, const int *__restrict__ col,
                                                     int upper, const int *__restrict__ depptr, int *fillpos,
                                                     int *__restrict__ deps) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        for (int q = row[i]; q < row[i + 1]; ++q) {
            const int d = col[q];
            if (depends(d, i, upper)) deps[depptr[d] + atomicAdd(&fillpos[d], 1)] = i;
        }
}

// Level 0: the rows that wait for nobody.
__global__ __launch_bounds__(kBlock) void k_frontier0(int n, const int *__restrict__ indeg, int *__restrict__ front,
                                                      int *cnt, int *__restrict__ level) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        if (indeg[i] == 0) {
            front[atomicAdd(&cnt[0], 1)] = i;
            level[i] = 0;
        }
}

// Level L: rows front[base[L] .. base[L] + cnt[L]).  COLOR: each takes the
// smallest colour >= 1 not held by a lower neighbour.  Then every row waiting
// for it counts down; those reaching zero are appended to level L + 1.
template <bool COLOR>
__global__ __launch_bounds__(kBlock) void k_level(int L, const int *__restrict__ row, const int *__restrict__ col,
                                                  const int *__restrict__ depptr, const int *__restrict__ deps,
                                                  int *indeg, int *front, int *base, int *cnt, int *level, int *clr) {
    const int b0 = base[L], sz = cnt[L], nb = b0 + sz;
    if (blockIdx.x == 0 && threadIdx.x == 0) base[L + 1] = nb;
    for (int t = blockIdx.x * kBlock + threadIdx.x; t < sz; t += gridDim.x * kBlock) {
        const int j = front[b0 + t];
        if (COLOR) {
            const int j0 = row[j], j1 = row[j + 1];
            unsigned long long held = 0;   // bit c-1: colour c held by a lower neighbour (c <= 64)
            for (int q = j0; q < j1; ++q) {
                const int d = col[q];
                if (d < j) {
                    const int c = clr[d];
                    if (c <= 64) held |= 1ull << (c - 1);
                }
            }
            int c;
            if (~held != 0ull) {
                c = __builtin_ctzll(~held) + 1;
            } else {   // colours 1..64 all held: the first free one above
                c = 65;
                while (true) {
                    bool taken = false;
                    for (int q = j0; q < j1 && !taken; ++q) taken = col[q] < j && clr[col[q]] == c;
                    if (!taken) break;
                    ++c;
                }
            }
            clr[j] = c;
        }
        for (int q = depptr[j]; q < depptr[j + 1]; ++q) {
            const int i = deps[q];
            if (atomicSub(&indeg[i], 1) == 1) {
                level[i] = L + 1;
                front[nb + atomicAdd(&cnt[L + 1], 1)] = i;
            }
        }
    }
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

// Levels (and first-fit colours) of all rows by the level-synchronous pass.
// On return *lev (and *clr) hold them; *has_off tells whether any row has an
// off-diagonal entry.
int topo(const hfh_csr *A, bool upper, bool color, Bufs &B, int **lev, int **clr, bool *has_off) {
    hfh_ctx *c = A->ctx;
    hipStream_t s = c->stream;
    const int n = A->nrow;
    int *indeg = nullptr, *depcnt = nullptr, *depptr = nullptr, *fillpos = nullptr, *deps = nullptr;
    int *front = nullptr, *base = nullptr, *cnt = nullptr, *level = nullptr, *colour = nullptr, *flags = nullptr;
    int rc;
    if ((rc = B.get(reinterpret_cast<void **>(&indeg), sizeof(int) * n)) ||
        (rc = B.get(reinterpret_cast<void **>(&depcnt), sizeof(int) * (n + 1))) ||
        (rc = B.get(reinterpret_cast<void **>(&depptr), sizeof(int) * (n + 1))) ||
        (rc = B.get(reinterpret_cast<void **>(&fillpos), sizeof(int) * n)) ||
        (rc = B.get(reinterpret_cast<void **>(&front), sizeof(int) * n)) ||
        (rc = B.get(reinterpret_cast<void **>(&base), sizeof(int) * (n + 2))) ||
        (rc = B.get(reinterpret_cast<void **>(&cnt), sizeof(int) * (n + 2))) ||
        (rc = B.get(reinterpret_cast<void **>(&level), sizeof(int) * n)) ||
        (rc = B.get(reinterpret_cast<void **>(&flags), sizeof(int))))
        return rc;
    if (color && (rc = B.get(reinterpret_cast<void **>(&colour), sizeof(int) * n))) return rc;
    const int g = grid_for(c, n);
    const int up = upper ? 1 : 0;
    HFH_HIP(hipMemsetAsync(depcnt, 0, sizeof(int) * (n + 1), s));
    HFH_HIP(hipMemsetAsync(fillpos, 0, sizeof(int) * n, s));
    HFH_HIP(hipMemsetAsync(base, 0, sizeof(int) * (n + 2), s));
    HFH_HIP(hipMemsetAsync(cnt, 0, sizeof(int) * (n + 2), s));
    HFH_HIP(hipMemsetAsync(flags, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_dep_count, dim3(g), dim3(kBlock), 0, s, n, A->row, A->col, up, indeg, depcnt, flags);
    if ((rc = launch_check("k_dep_count"))) return rc;
    size_t tmp_bytes = 0;
    HFH_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, depcnt, depptr, n + 1, s));
    void *tmp = nullptr;
    if ((rc = B.get(&tmp, std::max<size_t>(tmp_bytes, 1)))) return rc;
    HFH_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, depcnt, depptr, n + 1, s));
    int *h = reinterpret_cast<int *>(c->host_res);   // pinned, room for 256 doubles
    HFH_HIP(hipMemcpyAsync(h, depptr + n, sizeof(int), hipMemcpyDeviceToHost, s));
    HFH_HIP(hipStreamSynchronize(s));
    const long long ndeps = h[0];
    if ((rc = B.get(reinterpret_cast<void **>(&deps), sizeof(int) * static_cast<size_t>(std::max(ndeps, 1LL)))))
        return rc;
    hipLaunchKernelGGL(k_dep_fill, dim3(g), dim3(kBlock), 0, s, n, A->row, A->col, up, depptr, fillpos, deps);
    hipLaunchKernelGGL(k_frontier0, dim3(g), dim3(kBlock), 0, s, n, indeg, front, cnt, level);
    if ((rc = launch_check("k_frontier0"))) return rc;
    // every row joins a level after at most n levels (the dependencies
    // point one way, so they form no cycle); a frontier is at most a few
    // planes of a grid, so the level launches take a smaller grid
    const int gl = std::min(g, 512);
    for (int L = 0; L <= n; L += kChunk) {
        for (int k = 0; k < kChunk; ++k) {
            if (color)
                hipLaunchKernelGGL(k_level<true>, dim3(gl), dim3(kBlock), 0, s, L + k, A->row, A->col, depptr, deps,
                                   indeg, front, base, cnt, level, colour);
            else
                hipLaunchKernelGGL(k_level<false>, dim3(gl), dim3(kBlock), 0, s, L + k, A->row, A->col, depptr, deps,
                                   indeg, front, base, cnt, level, nullptr);
            if ((rc = launch_check("k_level"))) return rc;
            if (L + k + 1 > n) break;
        }
        const int next = std::min(L + kChunk, n + 1);
        HFH_HIP(hipMemcpyAsync(h, base + next, sizeof(int), hipMemcpyDeviceToHost, s));
        HFH_HIP(hipMemcpyAsync(h + 1, cnt + next, sizeof(int), hipMemcpyDeviceToHost, s));
        HFH_HIP(hipMemcpyAsync(h + 2, flags, sizeof(int), hipMemcpyDeviceToHost, s));
        HFH_HIP(hipStreamSynchronize(s));
        if (h[1] == 0) {
            HFH_CHECK(h[0] == n, HFH_ERR_INVALID, "device analysis: %d of %d rows placed", h[0], n);
            *lev = level;
            if (clr) *clr = colour;
            if (has_off) *has_off = h[2] != 0;
            return HFH_OK;
        }
    }
    set_error("device analysis: rows left after %d levels", n + 1);
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
    if ((rc = topo(A, upper != 0, false, B, &lev, nullptr, nullptr))) return rc;
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
    if ((rc = topo(A, false, true, B, &lev, &clr, &off))) return rc;
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
