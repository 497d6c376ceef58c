// hfh_spmv.hip -- CSR storage and the streaming CSR SpMV for gfx950.
//
// SpMV (lMatrix::VectorMult, CPUsimple_csr_lmatrix.cc:75-105) is HBM-bound:
// per call it must stream val (8 B/nnz) + col (4 B/nnz) + row (4 B/row) and
// write y (8 B/row); x is gathered through L2 / Infinity Cache.  The rows are
// partitioned into row blocks holding at most kSpmvTile nonzeros.  A
// workgroup streams its block's val/col with coalesced, non-temporal loads
// (so the streams do not evict the x window from L2), forms the products
// val[j]*x[col[j]] in LDS, and then one lane per row sums its products in
// ascending j.  The products are rounded before the sum exactly like the
// reference's `res += val[j]*x[col[j]]`, so rows that fit in one block agree
// bit for bit with the CPU path.  Rows longer than a block get a workgroup of
// their own and a tree sum.  Workgroups are remapped so each XCD streams one
// contiguous eighth of the matrix (its L2 then holds one x window).
#include "hfh_internal.h"

#include <algorithm>

namespace hfh {

template <typename T, bool ADD, bool DOT>
__global__ __launch_bounds__(kBlock) void k_spmv_stream(
    const int *__restrict__ row, const int *__restrict__ col, const T *__restrict__ val,
    const T *__restrict__ x, T *__restrict__ y, const int *__restrict__ rblk, int nblk,
    const T *__restrict__ w, double *part, unsigned int *counter, double *dst,
    const int *__restrict__ done) {
    if (done && *done) return;
    __shared__ T prod[kSpmvTile];
    __shared__ T sht[kBlock / 64];
    __shared__ double shd[kBlock / 64];
    const int b = xcd_remap(blockIdx.x, nblk);
    const int r0 = rblk[b], r1 = rblk[b + 1];
    const int j0 = row[r0], j1 = row[r1];
    const int cnt = j1 - j0;
    double dsum = 0;
    if (cnt <= kSpmvTile) {
#pragma unroll
        for (int k = 0; k < kSpmvTile / kBlock; ++k) {
            const int e = threadIdx.x + k * kBlock;
            if (e < cnt) {
                const int c = __builtin_nontemporal_load(&col[j0 + e]);
                const T v = __builtin_nontemporal_load(&val[j0 + e]);
                prod[e] = v * x[c];
            }
        }
        __syncthreads();
        for (int r = r0 + threadIdx.x; r < r1; r += kBlock) {
            const int a = row[r] - j0, e = row[r + 1] - j0;
            T s = 0;
            for (int j = a; j < e; ++j) s += prod[j];
            if (ADD) s = y[r] + s;
            y[r] = s;
            if (DOT) dsum += static_cast<double>(s) * static_cast<double>(w[r]);
        }
    } else {
        // one long row (r1 == r0 + 1): strided partial sums, fixed tree
        T s = 0;
        for (int j = j0 + threadIdx.x; j < j1; j += kBlock) s += val[j] * x[col[j]];
        s = block_sum(s, sht);
        if (threadIdx.x == 0) {
            if (ADD) s = y[r0] + s;
            y[r0] = s;
            if (DOT) dsum = static_cast<double>(s) * static_cast<double>(w[r0]);
        }
    }
    if (DOT) {
        __syncthreads();
        const double v[1] = {block_sum(dsum, shd)};
        finish_reduce<1>(v, part, 0, b, counter, nblk, dst);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_find_diag(int n, const int *__restrict__ row,
                                                      const int *__restrict__ col, int *__restrict__ diag) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    int p = -1;
    for (int j = row[i]; j < row[i + 1]; ++j)
        if (col[j] == i) p = j;
    diag[i] = p;
}

// extract_invdiagelements (lmatrix_csr_cpu.cc:2414-2432); rows without a
// diagonal entry keep 0 in the output (the reference leaves them unset).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_invdiag(int n, const int *__restrict__ diag,
                                                    const T *__restrict__ val, T *__restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int p = diag[i];
    out[i] = p >= 0 ? static_cast<T>(1) / val[p] : static_cast<T>(0);
}

template <typename TO, typename TI>
__global__ __launch_bounds__(kBlock) void k_cast_vals(long long n, TO *__restrict__ y, const TI *__restrict__ x) {
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        y[i] = static_cast<TO>(x[i]);
}

// Row blocks: consecutive rows while the block holds <= kSpmvTile nonzeros and
// <= kBlock rows; a longer row forms a block alone.
static std::vector<int> make_row_blocks(const std::vector<int> &row) {
    const int n = static_cast<int>(row.size()) - 1;
    std::vector<int> rb;
    rb.reserve(n / 16 + 2);
    rb.push_back(0);
    int r = 0;
    while (r < n) {
        const int start = r;
        if (row[r + 1] - row[r] > kSpmvTile) {
            r += 1;
        } else {
            while (r < n && r - start < kBlock && row[r + 1] - row[start] <= kSpmvTile) ++r;
        }
        rb.push_back(r);
    }
    return rb;
}

int spmv_launch_ex(const hfh_csr *A, const void *x, void *y, int mode, const void *w, double *part,
                   unsigned int *counter, double *dst, const int *done) {
    if (A->nrow == 0) return HFH_OK;
    hipStream_t s = A->ctx->stream;
    const dim3 g(A->nrblk), bl(kBlock);
    const bool dot = w != nullptr;
#define HFH_SPMV_LAUNCH(T, ADD, DOT)                                                                 \
    hipLaunchKernelGGL((k_spmv_stream<T, ADD, DOT>), g, bl, 0, s, A->row, A->col, (const T *)A->val, \
                       (const T *)x, (T *)y, A->rblk, A->nrblk, (const T *)w, part, counter, dst, done)
    if (A->dtype == HFH_F64) {
        if (mode == 0 && !dot) HFH_SPMV_LAUNCH(double, false, false);
        else if (mode == 0) HFH_SPMV_LAUNCH(double, false, true);
        else if (!dot) HFH_SPMV_LAUNCH(double, true, false);
        else HFH_SPMV_LAUNCH(double, true, true);
    } else {
        if (mode == 0 && !dot) HFH_SPMV_LAUNCH(float, false, false);
        else if (mode == 0) HFH_SPMV_LAUNCH(float, false, true);
        else if (!dot) HFH_SPMV_LAUNCH(float, true, false);
        else HFH_SPMV_LAUNCH(float, true, true);
    }
#undef HFH_SPMV_LAUNCH
    return launch_check("k_spmv_stream");
}

int spmv_launch(const hfh_csr *A, const void *x, void *y, int mode) {
    ScopedKernelTimer t(A->ctx, 0);
    return spmv_launch_ex(A, x, y, mode, nullptr, nullptr, nullptr, nullptr, nullptr);
}

}  // namespace hfh

using namespace hfh;

extern "C" int hfh_csr_create(hfh_ctx *ctx, int nrow, int ncol, int nnz, int dtype, hfh_csr **out) {
    HFH_CHECK(ctx && out, HFH_ERR_INVALID, "hfh_csr_create: null pointer");
    HFH_CHECK(nrow >= 0 && ncol >= 0 && nnz >= 0, HFH_ERR_INVALID, "hfh_csr_create: negative size");
    HFH_CHECK(dtype == HFH_F64 || dtype == HFH_F32, HFH_ERR_INVALID, "hfh_csr_create: bad dtype %d", dtype);
    hfh_csr *A = new hfh_csr();
    A->ctx = ctx;
    A->nrow = nrow;
    A->ncol = ncol;
    A->nnz = nnz;
    A->dtype = dtype;
    int rc;
    if ((rc = dev_alloc(reinterpret_cast<void **>(&A->row), sizeof(int) * (nrow + 1))) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&A->col), sizeof(int) * static_cast<size_t>(nnz))) ||
        (rc = dev_alloc(&A->val, A->elem() * static_cast<size_t>(nnz))) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&A->diag), sizeof(int) * static_cast<size_t>(nrow)))) {
        hfh_csr_destroy(A);
        return rc;
    }
    *out = A;
    return HFH_OK;
}

extern "C" int hfh_csr_destroy(hfh_csr *A) {
    if (!A) return HFH_OK;
    if (A->ctx && A->ctx->stream) (void)hipStreamSynchronize(A->ctx->stream);
    if (A->row) (void)hipFree(A->row);
    if (A->col) (void)hipFree(A->col);
    if (A->val) (void)hipFree(A->val);
    if (A->rblk) (void)hipFree(A->rblk);
    if (A->diag) (void)hipFree(A->diag);
    delete A;
    return HFH_OK;
}

extern "C" int hfh_csr_dims(const hfh_csr *A, int *nrow, int *ncol, int *nnz, int *dtype) {
    HFH_CHECK(A, HFH_ERR_INVALID, "hfh_csr_dims: null matrix");
    if (nrow) *nrow = A->nrow;
    if (ncol) *ncol = A->ncol;
    if (nnz) *nnz = A->nnz;
    if (dtype) *dtype = A->dtype;
    return HFH_OK;
}

extern "C" int hfh_csr_upload(hfh_csr *A, const int *row, const int *col, const void *val) {
    HFH_CHECK(A && row && (A->nnz == 0 || (col && val)), HFH_ERR_INVALID, "hfh_csr_upload: null pointer");
    const int n = A->nrow;
    HFH_CHECK(row[0] == 0, HFH_ERR_INVALID, "hfh_csr_upload: row[0] = %d, expected 0", row[0]);
    HFH_CHECK(row[n] == A->nnz, HFH_ERR_INVALID, "hfh_csr_upload: row[%d] = %d, expected nnz = %d", n, row[n], A->nnz);
    for (int i = 0; i < n; ++i) {
        HFH_CHECK(row[i + 1] >= row[i], HFH_ERR_INVALID, "hfh_csr_upload: row pointer decreases at row %d", i);
        for (int j = row[i]; j < row[i + 1]; ++j) {
            HFH_CHECK(col[j] >= 0 && col[j] < A->ncol, HFH_ERR_INVALID,
                      "hfh_csr_upload: col[%d] = %d outside [0, %d) in row %d", j, col[j], A->ncol, i);
            HFH_CHECK(j == row[i] || col[j] > col[j - 1], HFH_ERR_INVALID,
                      "hfh_csr_upload: columns of row %d not strictly ascending at entry %d", i, j);
        }
    }
    hipStream_t s = A->ctx->stream;
    A->h_row.assign(row, row + n + 1);
    std::vector<int> rb = make_row_blocks(A->h_row);
    if (A->rblk) (void)hipFree(A->rblk);
    A->rblk = nullptr;
    int rc = dev_alloc(reinterpret_cast<void **>(&A->rblk), sizeof(int) * rb.size());
    if (rc) return rc;
    A->nrblk = static_cast<int>(rb.size()) - 1;
    HFH_HIP(hipMemcpyAsync(A->row, row, sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    if (A->nnz) {
        HFH_HIP(hipMemcpyAsync(A->col, col, sizeof(int) * static_cast<size_t>(A->nnz), hipMemcpyHostToDevice, s));
        HFH_HIP(hipMemcpyAsync(A->val, val, A->elem() * static_cast<size_t>(A->nnz), hipMemcpyHostToDevice, s));
    }
    HFH_HIP(hipMemcpyAsync(A->rblk, rb.data(), sizeof(int) * rb.size(), hipMemcpyHostToDevice, s));
    if (n > 0) {
        hipLaunchKernelGGL(k_find_diag<int>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, A->row, A->col, A->diag);
        if ((rc = launch_check("k_find_diag"))) return rc;
    }
    HFH_HIP(hipStreamSynchronize(s));
    A->has_structure = true;
    return HFH_OK;
}

extern "C" int hfh_csr_upload_values(hfh_csr *A, const void *val) {
    HFH_CHECK(A && (val || A->nnz == 0), HFH_ERR_INVALID, "hfh_csr_upload_values: null pointer");
    HFH_CHECK(A->has_structure, HFH_ERR_INVALID, "hfh_csr_upload_values: structure not set");
    if (A->nnz == 0) return HFH_OK;
    HFH_HIP(hipMemcpyAsync(A->val, val, A->elem() * static_cast<size_t>(A->nnz), hipMemcpyHostToDevice, A->ctx->stream));
    HFH_HIP(hipStreamSynchronize(A->ctx->stream));
    return HFH_OK;
}

extern "C" int hfh_csr_download(const hfh_csr *A, int *row, int *col, void *val) {
    HFH_CHECK(A && A->has_structure, HFH_ERR_INVALID, "hfh_csr_download: no structure");
    hipStream_t s = A->ctx->stream;
    if (row) HFH_HIP(hipMemcpyAsync(row, A->row, sizeof(int) * (A->nrow + 1), hipMemcpyDeviceToHost, s));
    if (col && A->nnz) HFH_HIP(hipMemcpyAsync(col, A->col, sizeof(int) * static_cast<size_t>(A->nnz), hipMemcpyDeviceToHost, s));
    if (val && A->nnz) HFH_HIP(hipMemcpyAsync(val, A->val, A->elem() * static_cast<size_t>(A->nnz), hipMemcpyDeviceToHost, s));
    HFH_HIP(hipStreamSynchronize(s));
    return HFH_OK;
}

extern "C" int hfh_csr_cast(hfh_csr *dst, const hfh_csr *src) {
    HFH_CHECK(dst && src && src->has_structure, HFH_ERR_INVALID, "hfh_csr_cast: bad argument");
    HFH_CHECK(dst->nrow == src->nrow && dst->ncol == src->ncol && dst->nnz == src->nnz, HFH_ERR_INVALID,
              "hfh_csr_cast: dimension mismatch");
    hipStream_t s = dst->ctx->stream;
    HFH_HIP(hipMemcpyAsync(dst->row, src->row, sizeof(int) * (src->nrow + 1), hipMemcpyDeviceToDevice, s));
    if (src->nnz) HFH_HIP(hipMemcpyAsync(dst->col, src->col, sizeof(int) * static_cast<size_t>(src->nnz), hipMemcpyDeviceToDevice, s));
    if (src->nrow) HFH_HIP(hipMemcpyAsync(dst->diag, src->diag, sizeof(int) * static_cast<size_t>(src->nrow), hipMemcpyDeviceToDevice, s));
    dst->h_row = src->h_row;
    if (dst->rblk) (void)hipFree(dst->rblk);
    dst->rblk = nullptr;
    int rc = dev_alloc(reinterpret_cast<void **>(&dst->rblk), sizeof(int) * (src->nrblk + 1));
    if (rc) return rc;
    dst->nrblk = src->nrblk;
    HFH_HIP(hipMemcpyAsync(dst->rblk, src->rblk, sizeof(int) * (src->nrblk + 1), hipMemcpyDeviceToDevice, s));
    if (src->nnz) {
        const long long nnz = src->nnz;
        const int nb = static_cast<int>(std::min<long long>((nnz + kBlock - 1) / kBlock, 4096));
        if (dst->dtype == src->dtype)
            HFH_HIP(hipMemcpyAsync(dst->val, src->val, src->elem() * nnz, hipMemcpyDeviceToDevice, s));
        else if (dst->dtype == HFH_F32)
            hipLaunchKernelGGL((k_cast_vals<float, double>), dim3(nb), dim3(kBlock), 0, s, nnz, (float *)dst->val, (const double *)src->val);
        else
            hipLaunchKernelGGL((k_cast_vals<double, float>), dim3(nb), dim3(kBlock), 0, s, nnz, (double *)dst->val, (const float *)src->val);
        if ((rc = launch_check("k_cast_vals"))) return rc;
    }
    HFH_HIP(hipStreamSynchronize(s));
    dst->has_structure = true;
    return HFH_OK;
}

static int spmv_args(const hfh_csr *A, const hfh_vec *x, const hfh_vec *y, const char *what) {
    HFH_CHECK(A && x && y, HFH_ERR_INVALID, "%s: null argument", what);
    HFH_CHECK(A->has_structure, HFH_ERR_INVALID, "%s: matrix has no structure", what);
    HFH_CHECK(x->n == A->ncol && y->n == A->nrow, HFH_ERR_INVALID,
              "%s: size mismatch (A %dx%d, x %d, y %d)", what, A->nrow, A->ncol, x->n, y->n);
    HFH_CHECK(x->dtype == A->dtype && y->dtype == A->dtype, HFH_ERR_INVALID, "%s: dtype mismatch", what);
    HFH_CHECK(x->d != y->d, HFH_ERR_INVALID, "%s: x and y must not alias", what);
    return HFH_OK;
}

extern "C" int hfh_spmv(const hfh_csr *A, const hfh_vec *x, hfh_vec *y) {
    int rc = spmv_args(A, x, y, "hfh_spmv");
    if (rc) return rc;
    return spmv_launch(A, x->d, y->d, 0);
}

extern "C" int hfh_spmv_add(const hfh_csr *A, const hfh_vec *x, hfh_vec *y) {
    int rc = spmv_args(A, x, y, "hfh_spmv_add");
    if (rc) return rc;
    return spmv_launch(A, x->d, y->d, 1);
}

extern "C" int hfh_csr_invdiag(const hfh_csr *A, hfh_vec *invd) {
    HFH_CHECK(A && invd && A->has_structure, HFH_ERR_INVALID, "hfh_csr_invdiag: bad argument");
    HFH_CHECK(invd->n == A->nrow && invd->dtype == A->dtype, HFH_ERR_INVALID, "hfh_csr_invdiag: size/dtype mismatch");
    if (A->nrow == 0) return HFH_OK;
    const int nb = (A->nrow + kBlock - 1) / kBlock;
    if (A->dtype == HFH_F64)
        hipLaunchKernelGGL(k_invdiag<double>, dim3(nb), dim3(kBlock), 0, A->ctx->stream, A->nrow, A->diag, (const double *)A->val, (double *)invd->d);
    else
        hipLaunchKernelGGL(k_invdiag<float>, dim3(nb), dim3(kBlock), 0, A->ctx->stream, A->nrow, A->diag, (const float *)A->val, (float *)invd->d);
    return launch_check("k_invdiag");
}
