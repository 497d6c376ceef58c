// hfh_matops.hip -- lMatrix<T> value operations on the gfx950 platform
// (lmp/lmatrix.h:40-448): Zeros, Scale, ScaleOffdiag, ZeroRows, ZeroCols and
// extract_(inv)diagelements over a row range.  Each changes (or reads) the
// device CSR values in place with the reference's per-entry arithmetic; a
// change bumps the matrix's version so that the SpMV copies are refilled and
// the preconditioners that follow the live operator refresh their values.
#include "hfh_internal.h"

#include <algorithm>

namespace hfh {
namespace {

inline int grid_of(long long n) {
    return static_cast<int>(std::max<long long>(1, std::min<long long>((n + kBlock - 1) / kBlock, 8192)));
}

// CPU_CSR_lMatrix::Scale / ScaleOffdiag (lmatrix_csr_cpu.cc:206-225): v *= a
// (OFF: entries off the diagonal only).
template <typename T, bool OFF>
__global__ __launch_bounds__(kBlock) void k_csr_scale(int nrow, const int *__restrict__ row, const int *__restrict__ col,
                                                      T *__restrict__ val, T a) {
    for (long long r = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; r < nrow;
         r += static_cast<long long>(gridDim.x) * kBlock)
        for (int j = row[r]; j < row[r + 1]; ++j)
            if (!OFF || col[j] != r) val[j] = mul_rn(val[j], a);
}

// CPU_CSR_lMatrix::ZeroRows (lmatrix_csr_cpu.cc:1047-1068): every entry of a
// listed row becomes 0, its diagonal entry alpha; the pattern is kept.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_zero_rows(int m, const int *__restrict__ rows, const int *__restrict__ row,
                                                      const int *__restrict__ col, T *__restrict__ val, T alpha) {
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < m;
         i += static_cast<long long>(gridDim.x) * kBlock) {
        const int r = rows[i];
        for (int j = row[r]; j < row[r + 1]; ++j) val[j] = col[j] == r ? alpha : T(0);
    }
}

// CPU_CSR_lMatrix::ZeroCols (lmatrix_csr_cpu.cc:1071-1098): every entry in a
// listed column becomes 0, the diagonal ones alpha.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_zero_cols(int nrow, const int *__restrict__ row, const int *__restrict__ col,
                                                      const unsigned char *__restrict__ listed, T *__restrict__ val,
                                                      T alpha) {
    for (long long r = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; r < nrow;
         r += static_cast<long long>(gridDim.x) * kBlock)
        for (int j = row[r]; j < row[r + 1]; ++j)
            if (listed[col[j]]) val[j] = col[j] == r ? alpha : T(0);
}

// extract_diagelements / extract_invdiagelements (lmatrix_csr_cpu.cc:2395-2432)
// over rows [r0, r0 + m): out[i] = a_ii or 1 / a_ii; a row without a stored
// diagonal gives 0.
template <typename T, bool INV>
__global__ __launch_bounds__(kBlock) void k_diag_range(int r0, int m, const int *__restrict__ diag,
                                                       const T *__restrict__ val, T *__restrict__ out) {
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < m;
         i += static_cast<long long>(gridDim.x) * kBlock) {
        const int d = diag[r0 + i];
        out[i] = d < 0 ? T(0) : (INV ? T(1) / val[d] : val[d]);
    }
}

int values_changed(hfh_csr *A) {
    ++A->version;
    A->sell_dirty = true;
    int rc = spmv_sell_fill(A);
    if (rc) return rc;
    HFH_HIP(hipStreamSynchronize(A->ctx->stream));
    return HFH_OK;
}

int check_mat(const hfh_csr *A, const char *what) {
    HFH_CHECK(A && A->has_structure, HFH_ERR_INVALID, "%s: matrix has no structure", what);
    return HFH_OK;
}

// upload of a host row list, checked against the matrix's rows / columns
int upload_list(hfh_csr *A, const int *idx, int m, int bound, int **out, const char *what) {
    for (int i = 0; i < m; ++i)
        HFH_CHECK(idx[i] >= 0 && idx[i] < bound, HFH_ERR_INVALID, "%s: index[%d]=%d outside [0,%d)", what, i, idx[i],
                  bound);
    int rc = dev_alloc(reinterpret_cast<void **>(out), sizeof(int) * std::max(m, 1));
    if (rc) return rc;
    HFH_HIP(hipMemcpyAsync(*out, idx, sizeof(int) * m, hipMemcpyHostToDevice, A->ctx->stream));
    return HFH_OK;
}

}  // namespace
}  // namespace hfh

using namespace hfh;

extern "C" int hfh_csr_zeros(hfh_csr *A) {
    if (int rc = check_mat(A, "hfh_csr_zeros")) return rc;
    if (A->nnz == 0) return HFH_OK;
    HFH_HIP(hipMemsetAsync(A->val, 0, A->elem() * static_cast<size_t>(A->nnz), A->ctx->stream));
    return values_changed(A);
}

extern "C" int hfh_csr_scale(hfh_csr *A, double alpha, int offdiag_only) {
    if (int rc = check_mat(A, "hfh_csr_scale")) return rc;
    if (A->nnz == 0) return HFH_OK;
    const int nb = grid_of(A->nrow);
    hipStream_t s = A->ctx->stream;
    if (A->dtype == HFH_F64) {
        if (offdiag_only)
            hipLaunchKernelGGL((k_csr_scale<double, true>), dim3(nb), dim3(kBlock), 0, s, A->nrow, A->row, A->col, (double *)A->val, alpha);
        else
            hipLaunchKernelGGL((k_csr_scale<double, false>), dim3(nb), dim3(kBlock), 0, s, A->nrow, A->row, A->col, (double *)A->val, alpha);
    } else {
        const float a = static_cast<float>(alpha);
        if (offdiag_only)
            hipLaunchKernelGGL((k_csr_scale<float, true>), dim3(nb), dim3(kBlock), 0, s, A->nrow, A->row, A->col, (float *)A->val, a);
        else
            hipLaunchKernelGGL((k_csr_scale<float, false>), dim3(nb), dim3(kBlock), 0, s, A->nrow, A->row, A->col, (float *)A->val, a);
    }
    int rc = launch_check("k_csr_scale");
    return rc ? rc : values_changed(A);
}

extern "C" int hfh_csr_zero_rows(hfh_csr *A, const int *rows, int count, double alpha) {
    if (int rc = check_mat(A, "hfh_csr_zero_rows")) return rc;
    HFH_CHECK(count >= 0 && (count == 0 || rows), HFH_ERR_INVALID, "hfh_csr_zero_rows: bad argument");
    if (count == 0) return HFH_OK;
    int *d = nullptr;
    int rc = upload_list(A, rows, count, A->nrow, &d, "hfh_csr_zero_rows");
    if (rc) return rc;
    const int nb = grid_of(count);
    hipStream_t s = A->ctx->stream;
    if (A->dtype == HFH_F64)
        hipLaunchKernelGGL(k_zero_rows<double>, dim3(nb), dim3(kBlock), 0, s, count, d, A->row, A->col, (double *)A->val, alpha);
    else
        hipLaunchKernelGGL(k_zero_rows<float>, dim3(nb), dim3(kBlock), 0, s, count, d, A->row, A->col, (float *)A->val,
                           static_cast<float>(alpha));
    rc = launch_check("k_zero_rows");
    (void)hipStreamSynchronize(s);
    (void)hipFree(d);
    return rc ? rc : values_changed(A);
}

extern "C" int hfh_csr_zero_cols(hfh_csr *A, const int *cols, int count, double alpha) {
    if (int rc = check_mat(A, "hfh_csr_zero_cols")) return rc;
    HFH_CHECK(count > 0 && cols, HFH_ERR_INVALID, "hfh_csr_zero_cols: needs a non-empty column list");
    std::vector<unsigned char> listed(std::max(A->ncol, 1), 0);
    for (int i = 0; i < count; ++i) {
        HFH_CHECK(cols[i] >= 0 && cols[i] < A->ncol, HFH_ERR_INVALID, "hfh_csr_zero_cols: index[%d]=%d outside [0,%d)",
                  i, cols[i], A->ncol);
        listed[cols[i]] = 1;
    }
    if (A->nnz == 0) return HFH_OK;
    unsigned char *d = nullptr;
    int rc = dev_alloc(reinterpret_cast<void **>(&d), listed.size());
    if (rc) return rc;
    hipStream_t s = A->ctx->stream;
    if (hipMemcpyAsync(d, listed.data(), listed.size(), hipMemcpyHostToDevice, s) != hipSuccess) {
        (void)hipFree(d);
        return hip_fail(hipErrorUnknown, "hfh_csr_zero_cols: upload");
    }
    const int nb = grid_of(A->nrow);
    if (A->dtype == HFH_F64)
        hipLaunchKernelGGL(k_zero_cols<double>, dim3(nb), dim3(kBlock), 0, s, A->nrow, A->row, A->col, d, (double *)A->val, alpha);
    else
        hipLaunchKernelGGL(k_zero_cols<float>, dim3(nb), dim3(kBlock), 0, s, A->nrow, A->row, A->col, d, (float *)A->val,
                           static_cast<float>(alpha));
    rc = launch_check("k_zero_cols");
    (void)hipStreamSynchronize(s);
    (void)hipFree(d);
    return rc ? rc : values_changed(A);
}

extern "C" int hfh_csr_diag(const hfh_csr *A, int start, int end, int invert, hfh_vec *out) {
    if (int rc = check_mat(A, "hfh_csr_diag")) return rc;
    HFH_CHECK(out && start >= 0 && end <= A->nrow && start <= end, HFH_ERR_INVALID,
              "hfh_csr_diag: rows [%d, %d) outside [0, %d)", start, end, A->nrow);
    HFH_CHECK(out->n == end - start && out->dtype == A->dtype, HFH_ERR_INVALID,
              "hfh_csr_diag: output must hold %d values of the matrix's type", end - start);
    const int m = end - start;
    if (m == 0) return HFH_OK;
    const int nb = grid_of(m);
    hipStream_t s = A->ctx->stream;
    if (A->dtype == HFH_F64) {
        if (invert)
            hipLaunchKernelGGL((k_diag_range<double, true>), dim3(nb), dim3(kBlock), 0, s, start, m, A->diag, (const double *)A->val, (double *)out->d);
        else
            hipLaunchKernelGGL((k_diag_range<double, false>), dim3(nb), dim3(kBlock), 0, s, start, m, A->diag, (const double *)A->val, (double *)out->d);
    } else {
        if (invert)
            hipLaunchKernelGGL((k_diag_range<float, true>), dim3(nb), dim3(kBlock), 0, s, start, m, A->diag, (const float *)A->val, (float *)out->d);
        else
            hipLaunchKernelGGL((k_diag_range<float, false>), dim3(nb), dim3(kBlock), 0, s, start, m, A->diag, (const float *)A->val, (float *)out->d);
    }
    return launch_check("k_diag_range");
}

namespace hfh {
namespace {

// CPU_CSR_lMatrix::Pjacobi (lmatrix_csr_cpu.cc:920-948): out_i = in_i / a_ii
// (a division, as the reference); rows without a stored diagonal keep out_i.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pjacobi(int n, const int *__restrict__ diag, const T *__restrict__ val,
                                                    const T *__restrict__ in, T *__restrict__ out) {
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * kBlock) {
        const int d = diag[i];
        if (d >= 0) out[i] = in[i] / val[d];
    }
}

}  // namespace
}  // namespace hfh

extern "C" int hfh_csr_pjacobi(const hfh_csr *A, const hfh_vec *in, hfh_vec *out) {
    if (int rc = check_mat(A, "hfh_csr_pjacobi")) return rc;
    HFH_CHECK(in && out && A->nrow == A->ncol && in->n == A->nrow && out->n == A->nrow && in->dtype == A->dtype &&
                  out->dtype == A->dtype,
              HFH_ERR_INVALID, "hfh_csr_pjacobi: size/dtype mismatch");
    if (A->nrow == 0) return HFH_OK;
    const int nb = grid_of(A->nrow);
    hipStream_t s = A->ctx->stream;
    if (A->dtype == HFH_F64)
        hipLaunchKernelGGL(k_pjacobi<double>, dim3(nb), dim3(kBlock), 0, s, A->nrow, A->diag, (const double *)A->val,
                           (const double *)in->d, (double *)out->d);
    else
        hipLaunchKernelGGL(k_pjacobi<float>, dim3(nb), dim3(kBlock), 0, s, A->nrow, A->diag, (const float *)A->val,
                           (const float *)in->d, (float *)out->d);
    return launch_check("k_pjacobi");
}
