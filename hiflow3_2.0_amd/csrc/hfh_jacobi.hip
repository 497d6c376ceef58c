// hfh_jacobi.hip -- the stationary Jacobi solver / smoother on the device
// (Jacobi<LAD>, linear_solver/jacobi.cc:44-366).
//
// Prepare (jacobi.cc:44-67) swaps each row's diagonal entry to the front of
// the row in the operator itself (SwapDiagElementsToRowFront,
// lmatrix_csr_cpu.cc:477-495); hfh_csr_swap_diag_front does the same to a
// device matrix, and the SpMV of the swapped matrix sums its rows in the new
// stored order, as the reference's does.
//
// One step (DoNormalIteration / DoDampedIteration):
//   y = VectorMultNoDiag(x)            row sum without the diagonal, stored order
//   y = b - y;  y .*= inv_diag
//   x = y                              normal
//   x = (1 - w) x + w y                damped (x *= 1-w, then x += w y)
// is one fused kernel per step: the no-diagonal row sum over the matrix's SELL
// copy (lane = row; with the LDS-staged or dictionary-coded copy, the SpMV
// kernel's Jacobi form)
// and the update in its epilogue, written to a second buffer (every row reads
// the old x).  Products and sums are rounded
// separately, so every step is bit-identical to the CPU loop.  Matrices whose
// SpMV layout is not SELL-64 run a one-thread-per-row CSR step.
//
// Traffic of a step: the SELL copy (12 B per stored slot in FP64) plus x, b,
// inv_diag read and x written (32 B per row in FP64), an HBM-bound stream.
#include "hfh_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace hfh {
namespace {

template <typename T>
__global__ __launch_bounds__(kBlock) void k_swap_diag_front(int n, const int *__restrict__ row, int *__restrict__ col,
                                                            T *__restrict__ val) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int j0 = row[i], j1 = row[i + 1];
    for (int j = j0; j < j1; ++j)
        if (col[j] == i) {
            const T tv = val[j0];
            const int tc = col[j0];
            val[j0] = val[j];
            col[j0] = col[j];
            val[j] = tv;
            col[j] = tc;
            break;
        }
}

template <typename T> struct JPair;
template <> struct JPair<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct JPair<float> { typedef float type __attribute__((ext_vector_type(2))); };
typedef int jint_pair __attribute__((ext_vector_type(2)));

// Step over the SELL-64 copy.  PAIRS: slots 2q, 2q+1 of a lane at
// base + q*128 + 2*lane and, for an odd slice width L, the last slot at
// base + (L-1)*64 + lane; otherwise slot k at base + k*64 + lane.
template <typename T, bool DAMPED, bool PAIRS>
__global__ __launch_bounds__(kBlock) void k_jacobi_sell(int nrow, int nslice, const int *__restrict__ sptr,
                                                        const int *__restrict__ rlen, const int *__restrict__ scol,
                                                        const T *__restrict__ sval, const T *__restrict__ x,
                                                        T *__restrict__ xn, const T *__restrict__ b,
                                                        const T *__restrict__ invd, T w, T omw) {
    constexpr int WPB = kBlock / 64;
    constexpr int U = 4;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    for (int s = blockIdx.x * WPB + wave; s < nslice; s += gridDim.x * WPB) {
        const int base = sptr[s], L = (sptr[s + 1] - base) >> 6;
        const int r = s * 64 + lane;
        const int len = r < nrow ? rlen[r] : 0;
        T acc = 0;
        if (PAIRS) {
            typedef typename JPair<T>::type T2;
            const int Q = L >> 1;
            for (int q = 0; q < Q; ++q) {
                const int idx = base + q * 128 + 2 * lane;
                const jint_pair c = __builtin_nontemporal_load(reinterpret_cast<const jint_pair *>(&scol[idx]));
                const T2 v = __builtin_nontemporal_load(reinterpret_cast<const T2 *>(&sval[idx]));
                const T xa = x[c.x], xb = x[c.y];
                if (2 * q < len && c.x != r) acc = add_rn(acc, mul_rn(static_cast<T>(v.x), xa));
                if (2 * q + 1 < len && c.y != r) acc = add_rn(acc, mul_rn(static_cast<T>(v.y), xb));
            }
            if (L & 1) {
                const int idx = base + Q * 128 + lane;
                const int c = __builtin_nontemporal_load(&scol[idx]);
                const T v = __builtin_nontemporal_load(&sval[idx]);
                const T xv = x[c];
                if (2 * Q < len && c != r) acc = add_rn(acc, mul_rn(v, xv));
            }
        } else {
            for (int k = 0; k < L; k += U) {
                int c[U];
                T v[U], xv[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int idx = base + min(k + u, L - 1) * 64 + lane;
                    c[u] = __builtin_nontemporal_load(&scol[idx]);
                    v[u] = __builtin_nontemporal_load(&sval[idx]);
                }
#pragma unroll
                for (int u = 0; u < U; ++u) xv[u] = x[c[u]];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (k + u < len && c[u] != r) acc = add_rn(acc, mul_rn(v[u], xv[u]));
            }
        }
        if (r < nrow) xn[r] = jacobi_update<T, DAMPED>(acc, x[r], b[r], invd[r], w, omw);
    }
}

template <typename T, bool DAMPED>
__global__ __launch_bounds__(kBlock) void k_jacobi_csr(int n, const int *__restrict__ row, const int *__restrict__ col,
                                                       const T *__restrict__ val, const T *__restrict__ x,
                                                       T *__restrict__ xn, const T *__restrict__ b,
                                                       const T *__restrict__ invd, T w, T omw) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    T acc = 0;
    for (int j = row[i]; j < row[i + 1]; ++j)
        if (col[j] != i) acc = add_rn(acc, mul_rn(val[j], x[col[j]]));
    xn[i] = jacobi_update<T, DAMPED>(acc, x[i], b[i], invd[i], w, omw);
}

// VectorMultNoDiag (lmatrix_csr_cpu.cc:521-538): one thread per row, stored order.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_spmv_nodiag(int n, const int *__restrict__ row,
                                                        const int *__restrict__ col, const T *__restrict__ val,
                                                        const T *__restrict__ x, T *__restrict__ y) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    T acc = 0;
    for (int j = row[i]; j < row[i + 1]; ++j)
        if (col[j] != i) acc = add_rn(acc, mul_rn(val[j], x[col[j]]));
    y[i] = acc;
}

// r = b - r (ScaleAdd(-1, b): -r + b, exact negation, one rounding)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_minus(int n, T *__restrict__ r, const T *__restrict__ b) {
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        r[i] = add_rn(b[i], -r[i]);
}

int grid_of(hfh_ctx *c, long long items) {
    long long nb = (items + kBlock - 1) / kBlock;
    const long long cap = static_cast<long long>(c->num_cu > 0 ? c->num_cu : 256) * 16;
    return static_cast<int>(std::max<long long>(1, std::min(nb, cap)));
}

// Launches one step x -> xn.
template <typename T>
int jacobi_step(const hfh_csr *A, const T *x, T *xn, const T *b, const T *invd, bool damped, T w, T omw) {
    hfh_ctx *c = A->ctx;
    hipStream_t s = c->stream;
    const int n = A->nrow;
    if (A->spmv_algo == HFH_SPMV_SELL_LDS || A->spmv_algo == HFH_SPMV_SELL_DICT)
        return spmv_coded_jacobi<T>(A, x, xn, b, invd, damped, w, omw);
    const bool sell = A->spmv_algo == HFH_SPMV_SELL || A->spmv_algo == HFH_SPMV_SELL_PAIRS;
    if (sell) {
        const bool pairs = A->spmv_algo == HFH_SPMV_SELL_PAIRS;
        const int need = (A->nslice + kBlock / 64 - 1) / (kBlock / 64);
        const dim3 g(std::max(need, 1)), bl(kBlock);
        const T *sv = static_cast<const T *>(A->sl_val);
        if (pairs && damped)
            hipLaunchKernelGGL((k_jacobi_sell<T, true, true>), g, bl, 0, s, n, A->nslice, A->sl_ptr, A->sl_len,
                               A->sl_col, sv, x, xn, b, invd, w, omw);
        else if (pairs)
            hipLaunchKernelGGL((k_jacobi_sell<T, false, true>), g, bl, 0, s, n, A->nslice, A->sl_ptr, A->sl_len,
                               A->sl_col, sv, x, xn, b, invd, w, omw);
        else if (damped)
            hipLaunchKernelGGL((k_jacobi_sell<T, true, false>), g, bl, 0, s, n, A->nslice, A->sl_ptr, A->sl_len,
                               A->sl_col, sv, x, xn, b, invd, w, omw);
        else
            hipLaunchKernelGGL((k_jacobi_sell<T, false, false>), g, bl, 0, s, n, A->nslice, A->sl_ptr, A->sl_len,
                               A->sl_col, sv, x, xn, b, invd, w, omw);
        return launch_check("k_jacobi_sell");
    }
    const dim3 g((n + kBlock - 1) / kBlock), bl(kBlock);
    const T *v = static_cast<const T *>(A->val);
    if (damped)
        hipLaunchKernelGGL((k_jacobi_csr<T, true>), g, bl, 0, s, n, A->row, A->col, v, x, xn, b, invd, w, omw);
    else
        hipLaunchKernelGGL((k_jacobi_csr<T, false>), g, bl, 0, s, n, A->row, A->col, v, x, xn, b, invd, w, omw);
    return launch_check("k_jacobi_csr");
}

// ||b - A x|| with the full SpMV in stored order; the value is rounded to T
// as the reference's Norm2 returns it.
template <typename T>
int residual(const hfh_csr *A, const T *x, const T *b, T *r, double *out) {
    hfh_ctx *c = A->ctx;
    int rc = spmv_launch(A, x, r, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(k_b_minus<T>, dim3(grid_of(c, A->nrow)), dim3(kBlock), 0, c->stream, A->nrow, r, b);
    if ((rc = launch_check("k_b_minus"))) return rc;
    double *dst = c->red + kRedBlocks * 8;
    if ((rc = dot_async(c, A->dtype, A->nrow, r, r, dst))) return rc;
    double d = 0;
    if ((rc = fetch(c, dst, &d, 1))) return rc;
    *out = static_cast<double>(static_cast<T>(std::sqrt(static_cast<T>(d))));
    return HFH_OK;
}

template <typename T>
int jacobi_impl(const hfh_csr *A, const T *b, T *x, int mode, double w_d, int inner_iter, int smooth_iter,
                const hfh_control &ctl, hfh_solve_info *info, double *hist) {
    hfh_ctx *c = A->ctx;
    hipStream_t s = c->stream;
    const int n = A->nrow;
    const bool damped = mode == 1 || mode == 3;
    const T w = static_cast<T>(w_d), omw = static_cast<T>(1.0 - static_cast<double>(w));
    T *invd = nullptr, *y = nullptr, *r = nullptr;
    auto release = [&]() {
        if (invd) (void)hipFree(invd);
        if (y) (void)hipFree(y);
        if (r) (void)hipFree(r);
    };
    int rc;
    if ((rc = dev_alloc(reinterpret_cast<void **>(&invd), sizeof(T) * n)) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&y), sizeof(T) * n)) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&r), sizeof(T) * n))) {
        release();
        return rc;
    }
    // extract_invdiagelements leaves rows without a diagonal at 0
    if (hipMemsetAsync(invd, 0, sizeof(T) * n, s) != hipSuccess) {
        release();
        set_error("hfh_jacobi_solve: clearing the inverse diagonal failed");
        return HFH_ERR_HIP;
    }
    {
        hfh_vec vi;
        vi.ctx = c;
        vi.n = n;
        vi.dtype = A->dtype;
        vi.d = invd;
        if ((rc = hfh_csr_invdiag(A, &vi))) {
            release();
            return rc;
        }
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        release();
        set_error("hfh_jacobi_solve: event creation failed");
        return HFH_ERR_HIP;
    }
    (void)hipEventRecord(e0, s);
    // ping-pong: cur holds the iterate
    T *cur = x, *nxt = y;
    int iter = 0, state = HFH_ITERATE, nh = 0;
    double first = 0, res = 0;
    auto step = [&]() -> int {
        ++iter;
        int e = jacobi_step<T>(A, cur, nxt, b, invd, damped, w, omw);
        std::swap(cur, nxt);
        return e;
    };
    auto check = [&](double v) {
        if (iter == 0) first = v;
        if (v < ctl.atol) return static_cast<int>(HFH_SUCCESS_ABS);
        if (v < ctl.rtol * first) return static_cast<int>(HFH_SUCCESS_REL);
        if (v > ctl.dtol * first) return static_cast<int>(HFH_FAIL_DIV);
        if (iter >= ctl.maxits) return static_cast<int>(HFH_FAIL_MAXITS);
        return static_cast<int>(HFH_ITERATE);
    };
    if (mode >= 2) {
        for (int sm = 0; sm < smooth_iter && !rc; ++sm)
            for (int i = 0; i < inner_iter && !rc; ++i) rc = step();
    } else {
        rc = residual<T>(A, cur, b, r, &res);
        if (!rc) {
            state = check(res);
            if (hist) hist[nh++] = res;
        }
        while (!rc && state == HFH_ITERATE) {
            for (int i = 0; i < inner_iter && !rc; ++i) rc = step();
            if (rc) break;
            if ((rc = residual<T>(A, cur, b, r, &res))) break;
            if (hist) hist[nh++] = res;
            state = check(res);
        }
    }
    if (!rc && cur != x) {
        const hipError_t e = hipMemcpyAsync(x, cur, sizeof(T) * n, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) rc = hip_fail(e, "hfh_jacobi_solve: copy of the iterate");
    }
    (void)hipEventRecord(e1, s);
    if (!rc) {
        const hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = hip_fail(e, "hfh_jacobi_solve");
    }
    float ms = 0;
    if (!rc) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    release();
    if (rc) return rc;
    if (info) {
        info->iters = iter;
        info->state = state;
        info->res = res;
        info->res0 = first;
        info->time_ms = ms;
    }
    if (mode >= 2) return HFH_OK;
    if (state == HFH_SUCCESS_ABS || state == HFH_SUCCESS_REL) return HFH_OK;
    set_error("hfh_jacobi_solve: stopped without convergence (state %d)", state);
    return HFH_ERR_NOT_CONVERGED;
}

}  // namespace
}  // namespace hfh

using namespace hfh;

extern "C" int hfh_csr_swap_diag_front(hfh_csr *A) {
    HFH_CHECK(A && A->has_structure, HFH_ERR_INVALID, "hfh_csr_swap_diag_front: matrix has no structure");
    const int n = A->nrow;
    if (n == 0 || A->nnz == 0) return HFH_OK;
    hipStream_t s = A->ctx->stream;
    const int nb = (n + kBlock - 1) / kBlock;
    if (A->dtype == HFH_F64)
        hipLaunchKernelGGL(k_swap_diag_front<double>, dim3(nb), dim3(kBlock), 0, s, n, A->row, A->col,
                           (double *)A->val);
    else
        hipLaunchKernelGGL(k_swap_diag_front<float>, dim3(nb), dim3(kBlock), 0, s, n, A->row, A->col,
                           (float *)A->val);
    int rc = launch_check("k_swap_diag_front");
    if (rc) return rc;
    // the same swap on the host copy of the columns
    for (int i = 0; i < n; ++i) {
        const int j0 = A->h_row[i];
        for (int j = j0; j < A->h_row[i + 1]; ++j)
            if (A->h_col[j] == i) {
                std::swap(A->h_col[j0], A->h_col[j]);
                break;
            }
    }
    A->sorted = false;
    if ((rc = spmv_refresh_structure(A))) return rc;
    HFH_HIP(hipStreamSynchronize(s));
    return HFH_OK;
}

extern "C" int hfh_csr_is_sorted(const hfh_csr *A, int *sorted) {
    HFH_CHECK(A && sorted, HFH_ERR_INVALID, "hfh_csr_is_sorted: null argument");
    *sorted = A->sorted ? 1 : 0;
    return HFH_OK;
}

extern "C" int hfh_spmv_nodiag(const hfh_csr *A, const hfh_vec *x, hfh_vec *y) {
    HFH_CHECK(A && x && y && A->has_structure, HFH_ERR_INVALID, "hfh_spmv_nodiag: bad argument");
    HFH_CHECK(A->nrow == A->ncol && x->n == A->ncol && y->n == A->nrow && x->dtype == A->dtype &&
                  y->dtype == A->dtype && x->d != y->d,
              HFH_ERR_INVALID, "hfh_spmv_nodiag: size/dtype mismatch or aliasing");
    if (A->nrow == 0) return HFH_OK;
    hipStream_t s = A->ctx->stream;
    const int nb = (A->nrow + kBlock - 1) / kBlock;
    if (A->dtype == HFH_F64)
        hipLaunchKernelGGL(k_spmv_nodiag<double>, dim3(nb), dim3(kBlock), 0, s, A->nrow, A->row, A->col,
                           (const double *)A->val, (const double *)x->d, (double *)y->d);
    else
        hipLaunchKernelGGL(k_spmv_nodiag<float>, dim3(nb), dim3(kBlock), 0, s, A->nrow, A->row, A->col,
                           (const float *)A->val, (const float *)x->d, (float *)y->d);
    return launch_check("k_spmv_nodiag");
}

extern "C" int hfh_jacobi_solve(const hfh_csr *A, const hfh_vec *b, hfh_vec *x, int mode, double w, int inner_iter,
                                int smooth_iter, const hfh_control *ctl, hfh_solve_info *info, double *hist) {
    HFH_CHECK(A && b && x, HFH_ERR_INVALID, "hfh_jacobi_solve: null argument");
    HFH_CHECK(A->has_structure && A->nrow == A->ncol, HFH_ERR_INVALID,
              "hfh_jacobi_solve: matrix must be square with a structure");
    HFH_CHECK(b->n == A->nrow && x->n == A->nrow && b->dtype == A->dtype && x->dtype == A->dtype, HFH_ERR_INVALID,
              "hfh_jacobi_solve: vector size/dtype mismatch");
    HFH_CHECK(x->d != b->d, HFH_ERR_INVALID, "hfh_jacobi_solve: x and b must not alias");
    HFH_CHECK(mode >= 0 && mode <= 3, HFH_ERR_INVALID, "hfh_jacobi_solve: mode %d not in [0, 3]", mode);
    HFH_CHECK(inner_iter > 0 && smooth_iter > 0, HFH_ERR_INVALID,
              "hfh_jacobi_solve: inner_iter and smooth_iter must be positive");
    HFH_CHECK(!(mode == 1 || mode == 3) || w > 0, HFH_ERR_INVALID, "hfh_jacobi_solve: damping w must be > 0");
    HFH_CHECK(mode >= 2 || ctl, HFH_ERR_INVALID, "hfh_jacobi_solve: solve modes need a control");
    HFH_CHECK(mode >= 2 || (ctl->maxits > 0 && ctl->atol >= 0 && ctl->rtol >= 0 && ctl->dtol > 1), HFH_ERR_INVALID,
              "hfh_jacobi_solve: control needs maxits > 0, atol >= 0, rtol >= 0, dtol > 1");
    const hfh_control c = ctl ? *ctl : hfh_control{1, 0, 0, 2};
    if (A->nrow == 0) {
        if (info) *info = hfh_solve_info{0, HFH_SUCCESS_ABS, 0, 0, 0};
        return HFH_OK;
    }
    if (A->dtype == HFH_F64)
        return jacobi_impl<double>(A, static_cast<const double *>(b->d), static_cast<double *>(x->d), mode, w,
                                   inner_iter, smooth_iter, c, info, hist);
    return jacobi_impl<float>(A, static_cast<const float *>(b->d), static_cast<float *>(x->d), mode, w, inner_iter,
                              smooth_iter, c, info, hist);
}
