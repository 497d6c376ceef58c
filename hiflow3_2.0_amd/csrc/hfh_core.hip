// hfh_core.hip -- context, errors, vectors and BLAS-1 of the gfx950 platform.
//
// BLAS-1 kernels are HBM streams: 16 B per lane per load where alignment
// allows, grid-stride over a device-independent partition so that the
// reductions are deterministic (the same partials in the same order on every
// run and every device).  A one-workgroup finisher kernel sums the partials.
#include "hfh_internal.h"

#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>

namespace hfh {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

bool graphs_enabled() {
    static const bool on = [] {
        const char *e = std::getenv("HFH_GRAPHS");
        return !(e && std::strcmp(e, "0") == 0);
    }();
    return on;
}

int dev_alloc(void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
        *p = nullptr;
        return HFH_ERR_NOMEM;
    }
    return HFH_OK;
}

int launch_check(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, what);
    return HFH_OK;
}

ScopedKernelTimer::ScopedKernelTimer(hfh_ctx *c, int k) : ctx(c), kclass(k), on(c->timing.on) {
    if (on) (void)hipEventRecord(ctx->ev0, ctx->stream);
}

ScopedKernelTimer::~ScopedKernelTimer() {
    if (!on) return;
    (void)hipEventRecord(ctx->ev1, ctx->stream);
    (void)hipEventSynchronize(ctx->ev1);
    float ms = 0;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) {
        ctx->timing.ms[kclass] += ms;
        ctx->timing.launches[kclass] += 1;
    }
}

// ----------------------------------------------------------------------------
// reduction helpers

// Workgroups (= partials) of a dot over n elements: one per kBlock elements,
// up to kDotBlocks, or up to kDotBlocksBig when the operands exceed the
// 256 MiB Infinity Cache (HBM streams need more loads in flight: measured on
// gfx950, 16-byte non-temporal loads from 8192 workgroups read 6.6 TB/s,
// plain loads from 4096 workgroups 5.7 TB/s).  A single-quantity reduction
// stores its partials contiguously, and the context's partial region holds
// kRedBlocks * 8 of them.
constexpr int kDotBlocks = 4096;
constexpr int kDotBlocksBig = 8192;
constexpr double kMallBytes = 256.0 * 1024 * 1024;
constexpr int kStreamBlocksBig = 16384;   // workgroups of an element-wise stream beyond the cache
static_assert(kDotBlocksBig <= kRedBlocks * 8, "dot partials must fit the context's partial region");

static inline int red_blocks(int n, int cap) {
    long long nb = (static_cast<long long>(n) + kBlock - 1) / kBlock;
    if (nb < 1) nb = 1;
    if (nb > cap) nb = cap;
    return static_cast<int>(nb);
}

// Sum of x_i y_i in FP64 (FP32 products are exact in FP64).  nv: 16-byte
// vectors read as such (0 when an operand is not 16-byte aligned); the
// elements from nv*W on are read one by one.  Each thread keeps kU vectors of
// both operands in flight and kU partial sums.  NT: non-temporal vector
// loads (operands that stream from HBM and would not stay in the cache).
constexpr int kU = 4;

template <typename T, bool NT>
__global__ __launch_bounds__(kBlock) void k_dot(int n, long long nv, const T *__restrict__ x,
                                                const T *__restrict__ y, double *part) {
    typedef typename V16<T>::type V;
    constexpr int W = V16<T>::W;
    __shared__ double sh[kBlock / 64];
    const V *__restrict__ xv = reinterpret_cast<const V *>(x);
    const V *__restrict__ yv = reinterpret_cast<const V *>(y);
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x;
    double acc[kU] = {0, 0, 0, 0};
    for (; i + (kU - 1) * stride < nv; i += kU * stride) {
        V a[kU], b[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (NT) {
                a[u] = __builtin_nontemporal_load(&xv[i + u * stride]);
                b[u] = __builtin_nontemporal_load(&yv[i + u * stride]);
            } else {
                a[u] = xv[i + u * stride];
                b[u] = yv[i + u * stride];
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
#pragma unroll
            for (int w = 0; w < W; ++w) acc[u] += static_cast<double>(a[u][w]) * static_cast<double>(b[u][w]);
    }
    for (; i < nv; i += stride) {
        const V a = xv[i], b = yv[i];
#pragma unroll
        for (int w = 0; w < W; ++w) acc[0] += static_cast<double>(a[w]) * static_cast<double>(b[w]);
    }
    for (long long j = nv * W + blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < n; j += stride)
        acc[1] += static_cast<double>(x[j]) * static_cast<double>(y[j]);
    const double v[1] = {block_sum((acc[0] + acc[1]) + (acc[2] + acc[3]), sh)};
    store_partials<1>(v, part, kRedBlocks, blockIdx.x);
}

// 16-byte vectors of both operands when both are aligned, else none.
template <typename T>
static inline long long vec_count(int n, const void *a, const void *b) {
    const bool aligned = ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) == 0;
    return aligned ? n / V16<T>::W : 0;
}

__global__ __launch_bounds__(kFinBlock) void k_fin_store(const double *part, int nb, double *dst) {
    double out[1];
    reduce_partials<1>(part, kRedBlocks, nb, out);
    if (threadIdx.x == 0) dst[0] = out[0];
}

// FP32 vectors: products formed exactly in FP64 and summed in FP64.
int dot_async(hfh_ctx *ctx, int dtype, int n, const void *x, const void *y, double *dst) {
    const bool big = 2.0 * n * (dtype == HFH_F64 ? 8 : 4) > kMallBytes;
    const int nb = red_blocks(n, big ? kDotBlocksBig : kDotBlocks);
    if (dtype == HFH_F64 && big)
        hipLaunchKernelGGL((k_dot<double, true>), dim3(nb), dim3(kBlock), 0, ctx->stream, n,
                           vec_count<double>(n, x, y), (const double *)x, (const double *)y, ctx->red);
    else if (dtype == HFH_F64)
        hipLaunchKernelGGL((k_dot<double, false>), dim3(nb), dim3(kBlock), 0, ctx->stream, n,
                           vec_count<double>(n, x, y), (const double *)x, (const double *)y, ctx->red);
    else if (big)
        hipLaunchKernelGGL((k_dot<float, true>), dim3(nb), dim3(kBlock), 0, ctx->stream, n, vec_count<float>(n, x, y),
                           (const float *)x, (const float *)y, ctx->red);
    else
        hipLaunchKernelGGL((k_dot<float, false>), dim3(nb), dim3(kBlock), 0, ctx->stream, n, vec_count<float>(n, x, y),
                           (const float *)x, (const float *)y, ctx->red);
    int rc = launch_check("k_dot");
    if (rc) return rc;
    hipLaunchKernelGGL(k_fin_store, dim3(1), dim3(kFinBlock), 0, ctx->stream, (const double *)ctx->red, nb, dst);
    return launch_check("k_fin_store");
}

int partials_sum(hfh_ctx *ctx, int nb, double *dst) {
    hipLaunchKernelGGL(k_fin_store, dim3(1), dim3(kFinBlock), 0, ctx->stream, (const double *)ctx->red, nb, dst);
    return launch_check("k_fin_store");
}

int fetch(hfh_ctx *ctx, const double *dsrc, double *host, int count) {
    HFH_HIP(hipMemcpyAsync(ctx->host_res, dsrc, sizeof(double) * count, hipMemcpyDeviceToHost,
                           ctx->stream));
    HFH_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < count; ++i) host[i] = ctx->host_res[i];
    return HFH_OK;
}

// ----------------------------------------------------------------------------
// element-wise kernels

// Element-wise updates y_i = op(y_i, x_i), every product and sum rounded
// separately (the reference's CPU loops, bit for bit):
//   Axpy      y + a x        (CPUsimple_lvector.cc:234-287)
//   ScaleAdd  a y + x        (CPUsimple_lvector.cc:310-365)
//   Scale     y a            (CPUsimple_lvector.cc:388-435)
//   EWMult    y x            (lvector_cpu.cc:571-578)
// 16-byte vectors, kU per thread in flight; nv as for k_dot.
template <typename T> struct OpAxpy {
    T a;
    __device__ __forceinline__ T operator()(T y, T x) const { return add_rn(y, mul_rn(a, x)); }
};
template <typename T> struct OpScaleAdd {
    T a;
    __device__ __forceinline__ T operator()(T y, T x) const { return add_rn(mul_rn(a, y), x); }
};
template <typename T> struct OpScale {
    T a;
    __device__ __forceinline__ T operator()(T y, T) const { return mul_rn(y, a); }
};
template <typename T> struct OpMult {
    __device__ __forceinline__ T operator()(T y, T x) const { return mul_rn(y, x); }
};

// the scalar of an op, rounded to the vector's type as the reference's
// lVector<T> interface receives it
template <typename T> inline void set_scalar(OpAxpy<T> &o, double a) { o.a = static_cast<T>(a); }
template <typename T> inline void set_scalar(OpScaleAdd<T> &o, double a) { o.a = static_cast<T>(a); }
template <typename T> inline void set_scalar(OpScale<T> &o, double a) { o.a = static_cast<T>(a); }
template <typename T> inline void set_scalar(OpMult<T> &, double) {}

// NT: non-temporal 16-byte loads and stores, for operands larger than the
// Infinity Cache (2 GiB FP64 vectors on gfx950: copy-like 4.52 -> 5.09 TB/s,
// triad-like 4.62 -> 4.81 TB/s with 16,384 instead of 4,096 workgroups).
template <typename T, typename Op, bool BIN, bool NT = false>
__global__ __launch_bounds__(kBlock) void k_update(int n, long long nv, T *__restrict__ y, const T *__restrict__ x,
                                                   Op op) {
    typedef typename V16<T>::type V;
    constexpr int W = V16<T>::W;
    V *__restrict__ yv = reinterpret_cast<V *>(y);
    const V *__restrict__ xv = reinterpret_cast<const V *>(x);
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x;
    for (; i + (kU - 1) * stride < nv; i += kU * stride) {
        V a[kU], b[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (NT) {
                a[u] = __builtin_nontemporal_load(&yv[i + u * stride]);
                if (BIN) b[u] = __builtin_nontemporal_load(&xv[i + u * stride]);
            } else {
                a[u] = yv[i + u * stride];
                if (BIN) b[u] = xv[i + u * stride];
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
#pragma unroll
            for (int w = 0; w < W; ++w) a[u][w] = op(a[u][w], BIN ? b[u][w] : T(0));
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (NT) __builtin_nontemporal_store(a[u], &yv[i + u * stride]);
            else yv[i + u * stride] = a[u];
        }
    }
    for (; i < nv; i += stride) {
        V a = yv[i], b;
        if (BIN) b = xv[i];
#pragma unroll
        for (int w = 0; w < W; ++w) a[w] = op(a[w], BIN ? b[w] : T(0));
        yv[i] = a;
    }
    for (long long j = nv * W + blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < n; j += stride)
        y[j] = op(y[j], BIN ? x[j] : T(0));
}

template <typename TO, typename TI>
__global__ __launch_bounds__(kBlock) void k_cast(int n, TO *__restrict__ y, const TI *__restrict__ x) {
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        y[i] = static_cast<TO>(x[i]);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather(int n, T *__restrict__ y, const T *__restrict__ x,
                                                   const int *__restrict__ idx) {
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        y[i] = x[idx[i]];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scatter(int n, T *__restrict__ y, const T *__restrict__ x,
                                                    const int *__restrict__ idx) {
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        y[idx[i]] = x[i];
}

static inline int ew_blocks(hfh_ctx *ctx, int n) {
    long long nb = (static_cast<long long>(n) + kBlock - 1) / kBlock;
    const long long cap = static_cast<long long>(ctx->num_cu > 0 ? ctx->num_cu : 256) * 16;
    if (nb > cap) nb = cap;
    if (nb < 1) nb = 1;
    return static_cast<int>(nb);
}

}  // namespace hfh

using namespace hfh;

// ============================================================================
// C ABI

extern "C" const char *hfh_last_error(void) { return hfh::g_err; }
extern "C" const char *hfh_version(void) { return "hfh-gfx950 0.1"; }

extern "C" int hfh_device_count(int *count) {
    HFH_CHECK(count, HFH_ERR_INVALID, "hfh_device_count: null pointer");
    HFH_HIP(hipGetDeviceCount(count));
    return HFH_OK;
}

extern "C" int hfh_ctx_create(int device, hfh_ctx **out) {
    HFH_CHECK(out, HFH_ERR_INVALID, "hfh_ctx_create: null pointer");
    *out = nullptr;
    int count = 0;
    HFH_HIP(hipGetDeviceCount(&count));
    HFH_CHECK(device >= 0 && device < count, HFH_ERR_INVALID,
              "hfh_ctx_create: device %d out of range (%d devices)", device, count);
    HFH_HIP(hipSetDevice(device));
    hfh_ctx *c = new hfh_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    HFH_HIP(hipGetDeviceProperties(&prop, device));
    c->num_cu = prop.multiProcessorCount;
    c->hbm = static_cast<long long>(prop.totalGlobalMem);
    snprintf(c->name, sizeof c->name, "%s", prop.gcnArchName);
    HFH_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    int rc = dev_alloc(reinterpret_cast<void **>(&c->red), sizeof(double) * (kRedBlocks * 8 + 256));
    if (rc) { delete c; return rc; }
    HFH_HIP(hipMemset(c->red, 0, sizeof(double) * (kRedBlocks * 8 + 256)));
    HFH_HIP(hipHostMalloc(reinterpret_cast<void **>(&c->host_res), sizeof(double) * 256));
    HFH_HIP(hipEventCreate(&c->ev0));
    HFH_HIP(hipEventCreate(&c->ev1));
    // the allocation-free CG's static array on this device (current now)
    if ((rc = static_scratch(device, &c->scratch))) {
        (void)hipStreamDestroy(c->stream);
        (void)hipFree(c->red);
        (void)hipHostFree(c->host_res);
        (void)hipEventDestroy(c->ev0);
        (void)hipEventDestroy(c->ev1);
        delete c;
        return rc;
    }
    *out = c;
    return HFH_OK;
}

extern "C" int hfh_ctx_destroy(hfh_ctx *c) {
    if (!c) return HFH_OK;
    HFH_CHECK(c->live == 0, HFH_ERR_INVALID, "hfh_ctx_destroy: %d objects still use the context; destroy them first",
              c->live);
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->red) (void)hipFree(c->red);
    if (c->host_res) (void)hipHostFree(c->host_res);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return HFH_OK;
}

extern "C" int hfh_ctx_sync(hfh_ctx *c) {
    HFH_CHECK(c, HFH_ERR_INVALID, "hfh_ctx_sync: null context");
    HFH_HIP(hipStreamSynchronize(c->stream));
    return HFH_OK;
}

extern "C" int hfh_ctx_info(hfh_ctx *c, int *num_cu, long long *hbm, char *name, int len) {
    HFH_CHECK(c, HFH_ERR_INVALID, "hfh_ctx_info: null context");
    if (num_cu) *num_cu = c->num_cu;
    if (hbm) *hbm = c->hbm;
    if (name && len > 0) snprintf(name, len, "%s", c->name);
    return HFH_OK;
}

extern "C" int hfh_ctx_set_timing(hfh_ctx *c, int on) {
    HFH_CHECK(c, HFH_ERR_INVALID, "hfh_ctx_set_timing: null context");
    c->timing.on = on != 0;
    return HFH_OK;
}

extern "C" int hfh_ctx_timing(hfh_ctx *c, int k, long long *launches, double *ms) {
    HFH_CHECK(c && k >= 0 && k < 4, HFH_ERR_INVALID, "hfh_ctx_timing: bad argument");
    if (launches) *launches = c->timing.launches[k];
    if (ms) *ms = c->timing.ms[k];
    return HFH_OK;
}

extern "C" int hfh_ctx_timing_reset(hfh_ctx *c) {
    HFH_CHECK(c, HFH_ERR_INVALID, "hfh_ctx_timing_reset: null context");
    for (int k = 0; k < 4; ++k) {
        c->timing.launches[k] = 0;
        c->timing.ms[k] = 0;
    }
    return HFH_OK;
}

// ---------------------------------------------------------------- vectors

extern "C" int hfh_vec_create(hfh_ctx *ctx, int n, int dtype, hfh_vec **out) {
    HFH_CHECK(ctx && out, HFH_ERR_INVALID, "hfh_vec_create: null pointer");
    HFH_CHECK(n >= 0, HFH_ERR_INVALID, "hfh_vec_create: negative size %d", n);
    HFH_CHECK(dtype == HFH_F64 || dtype == HFH_F32, HFH_ERR_INVALID, "hfh_vec_create: bad dtype %d", dtype);
    hfh_vec *v = new hfh_vec();
    v->ctx = ctx;
    v->n = n;
    v->dtype = dtype;
    int rc = dev_alloc(&v->d, v->elem() * static_cast<size_t>(n));
    if (rc) { delete v; return rc; }
    if (n > 0) {
        hipError_t e = hipMemsetAsync(v->d, 0, v->elem() * n, ctx->stream);
        if (e != hipSuccess) { (void)hipFree(v->d); delete v; return hip_fail(e, "hipMemsetAsync"); }
    }
    ++ctx->live;
    *out = v;
    return HFH_OK;
}

extern "C" int hfh_vec_destroy(hfh_vec *v) {
    if (!v) return HFH_OK;
    (void)hipStreamSynchronize(v->ctx->stream);
    if (v->d) (void)hipFree(v->d);
    if (v->iset) (void)hipFree(v->iset);
    --v->ctx->live;
    delete v;
    return HFH_OK;
}

extern "C" int hfh_vec_size(const hfh_vec *v, int *n) {
    HFH_CHECK(v && n, HFH_ERR_INVALID, "hfh_vec_size: null pointer");
    *n = v->n;
    return HFH_OK;
}

extern "C" int hfh_vec_dtype(const hfh_vec *v, int *dtype) {
    HFH_CHECK(v && dtype, HFH_ERR_INVALID, "hfh_vec_dtype: null pointer");
    *dtype = v->dtype;
    return HFH_OK;
}

extern "C" int hfh_vec_upload(hfh_vec *v, const void *host, int off, int cnt) {
    HFH_CHECK(v && (host || cnt == 0), HFH_ERR_INVALID, "hfh_vec_upload: null pointer");
    HFH_CHECK(off >= 0 && cnt >= 0 && static_cast<long long>(off) + cnt <= v->n, HFH_ERR_INVALID,
              "hfh_vec_upload: range [%d, %d) outside vector of size %d", off, off + cnt, v->n);
    if (cnt == 0) return HFH_OK;
    HFH_HIP(hipMemcpyAsync(static_cast<char *>(v->d) + v->elem() * off, host, v->elem() * cnt,
                           hipMemcpyHostToDevice, v->ctx->stream));
    HFH_HIP(hipStreamSynchronize(v->ctx->stream));
    return HFH_OK;
}

extern "C" int hfh_vec_download(const hfh_vec *v, void *host, int off, int cnt) {
    HFH_CHECK(v && (host || cnt == 0), HFH_ERR_INVALID, "hfh_vec_download: null pointer");
    HFH_CHECK(off >= 0 && cnt >= 0 && static_cast<long long>(off) + cnt <= v->n, HFH_ERR_INVALID,
              "hfh_vec_download: range [%d, %d) outside vector of size %d", off, off + cnt, v->n);
    if (cnt == 0) return HFH_OK;
    HFH_HIP(hipMemcpyAsync(host, static_cast<const char *>(v->d) + v->elem() * off, v->elem() * cnt,
                           hipMemcpyDeviceToHost, v->ctx->stream));
    HFH_HIP(hipStreamSynchronize(v->ctx->stream));
    return HFH_OK;
}

static int check_index(const hfh_vec *v, const int *index, int count, const char *what) {
    for (int i = 0; i < count; ++i)
        HFH_CHECK(index[i] >= 0 && index[i] < v->n, HFH_ERR_INVALID, "%s: index[%d]=%d outside [0,%d)",
                  what, i, index[i], v->n);
    return HFH_OK;
}

extern "C" int hfh_vec_get_values(const hfh_vec *v, const int *index, int count, void *values) {
    HFH_CHECK(v && count >= 0 && (count == 0 || (index && values)), HFH_ERR_INVALID,
              "hfh_vec_get_values: bad argument");
    if (count == 0) return HFH_OK;
    int rc = check_index(v, index, count, "hfh_vec_get_values");
    if (rc) return rc;
    void *dv = nullptr, *di = nullptr;
    if ((rc = dev_alloc(&dv, v->elem() * count))) return rc;
    if ((rc = dev_alloc(&di, sizeof(int) * count))) { (void)hipFree(dv); return rc; }
    hipStream_t s = v->ctx->stream;
    HFH_HIP(hipMemcpyAsync(di, index, sizeof(int) * count, hipMemcpyHostToDevice, s));
    const int nb = ew_blocks(v->ctx, count);
    if (v->dtype == HFH_F64)
        hipLaunchKernelGGL(k_gather<double>, dim3(nb), dim3(kBlock), 0, s, count, (double *)dv, (const double *)v->d, (const int *)di);
    else
        hipLaunchKernelGGL(k_gather<float>, dim3(nb), dim3(kBlock), 0, s, count, (float *)dv, (const float *)v->d, (const int *)di);
    rc = launch_check("k_gather");
    if (rc == HFH_OK) {
        HFH_HIP(hipMemcpyAsync(values, dv, v->elem() * count, hipMemcpyDeviceToHost, s));
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(dv);
    (void)hipFree(di);
    return rc;
}

extern "C" int hfh_vec_zeros(hfh_vec *v) {
    HFH_CHECK(v, HFH_ERR_INVALID, "hfh_vec_zeros: null vector");
    if (v->n == 0) return HFH_OK;
    HFH_HIP(hipMemsetAsync(v->d, 0, v->elem() * v->n, v->ctx->stream));
    return HFH_OK;
}

extern "C" int hfh_vec_copy(hfh_vec *dst, const hfh_vec *src) {
    HFH_CHECK(dst && src, HFH_ERR_INVALID, "hfh_vec_copy: null vector");
    HFH_CHECK(dst->n == src->n && dst->dtype == src->dtype, HFH_ERR_INVALID,
              "hfh_vec_copy: size/dtype mismatch (%d vs %d)", dst->n, src->n);
    if (dst == src || dst->n == 0) return HFH_OK;
    HFH_HIP(hipMemcpyAsync(dst->d, src->d, dst->elem() * dst->n, hipMemcpyDeviceToDevice, dst->ctx->stream));
    return HFH_OK;
}

extern "C" int hfh_vec_copy_range(hfh_vec *dst, int doff, const hfh_vec *src, int soff, int cnt) {
    HFH_CHECK(dst && src, HFH_ERR_INVALID, "hfh_vec_copy_range: null vector");
    HFH_CHECK(dst->dtype == src->dtype, HFH_ERR_INVALID, "hfh_vec_copy_range: dtype mismatch");
    HFH_CHECK(doff >= 0 && soff >= 0 && cnt >= 0 && (long long)doff + cnt <= dst->n &&
                  (long long)soff + cnt <= src->n,
              HFH_ERR_INVALID, "hfh_vec_copy_range: range out of bounds");
    if (cnt == 0) return HFH_OK;
    HFH_HIP(hipMemcpyAsync(static_cast<char *>(dst->d) + dst->elem() * doff,
                           static_cast<const char *>(src->d) + src->elem() * soff, dst->elem() * cnt,
                           hipMemcpyDeviceToDevice, dst->ctx->stream));
    return HFH_OK;
}

extern "C" int hfh_vec_cast(hfh_vec *dst, const hfh_vec *src) {
    HFH_CHECK(dst && src, HFH_ERR_INVALID, "hfh_vec_cast: null vector");
    HFH_CHECK(dst->n == src->n, HFH_ERR_INVALID, "hfh_vec_cast: size mismatch (%d vs %d)", dst->n, src->n);
    if (dst->n == 0) return HFH_OK;
    if (dst->dtype == src->dtype) return hfh_vec_copy(dst, src);
    const int nb = ew_blocks(dst->ctx, dst->n);
    hipStream_t s = dst->ctx->stream;
    if (dst->dtype == HFH_F32)
        hipLaunchKernelGGL((k_cast<float, double>), dim3(nb), dim3(kBlock), 0, s, dst->n, (float *)dst->d, (const double *)src->d);
    else
        hipLaunchKernelGGL((k_cast<double, float>), dim3(nb), dim3(kBlock), 0, s, dst->n, (double *)dst->d, (const float *)src->d);
    return launch_check("k_cast");
}

extern "C" int hfh_vec_reorder(hfh_vec *dst, const hfh_vec *src, const int *perm) {
    HFH_CHECK(dst && src && (perm || dst->n == 0), HFH_ERR_INVALID, "hfh_vec_reorder: null pointer");
    HFH_CHECK(dst != src, HFH_ERR_INVALID, "hfh_vec_reorder: in-place reorder not supported");
    HFH_CHECK(dst->n == src->n && dst->dtype == src->dtype, HFH_ERR_INVALID, "hfh_vec_reorder: mismatch");
    if (dst->n == 0) return HFH_OK;
    for (int i = 0; i < dst->n; ++i)
        HFH_CHECK(perm[i] >= 0 && perm[i] < src->n, HFH_ERR_INVALID, "hfh_vec_reorder: perm[%d]=%d out of range", i, perm[i]);
    void *di = nullptr;
    int rc = dev_alloc(&di, sizeof(int) * dst->n);
    if (rc) return rc;
    hipStream_t s = dst->ctx->stream;
    HFH_HIP(hipMemcpyAsync(di, perm, sizeof(int) * dst->n, hipMemcpyHostToDevice, s));
    const int nb = ew_blocks(dst->ctx, dst->n);
    if (dst->dtype == HFH_F64)
        hipLaunchKernelGGL(k_gather<double>, dim3(nb), dim3(kBlock), 0, s, dst->n, (double *)dst->d, (const double *)src->d, (const int *)di);
    else
        hipLaunchKernelGGL(k_gather<float>, dim3(nb), dim3(kBlock), 0, s, dst->n, (float *)dst->d, (const float *)src->d, (const int *)di);
    rc = launch_check("k_gather");
    (void)hipStreamSynchronize(s);
    (void)hipFree(di);
    return rc;
}

// ---------------------------------------------------------------- BLAS-1

static int same_shape(const hfh_vec *a, const hfh_vec *b, const char *what) {
    HFH_CHECK(a && b, HFH_ERR_INVALID, "%s: null vector", what);
    HFH_CHECK(a->n == b->n && a->dtype == b->dtype && a->ctx == b->ctx, HFH_ERR_INVALID,
              "%s: vectors differ in size/dtype/context (%d vs %d)", what, a->n, b->n);
    return HFH_OK;
}

extern "C" int hfh_dot(const hfh_vec *x, const hfh_vec *y, double *out) {
    int rc = same_shape(x, y, "hfh_dot");
    if (rc) return rc;
    HFH_CHECK(out, HFH_ERR_INVALID, "hfh_dot: null output");
    hfh_ctx *c = x->ctx;
    if (x->n == 0) { *out = 0; return HFH_OK; }
    double *dst = c->red + kRedBlocks * 8;
    if ((rc = dot_async(c, x->dtype, x->n, x->d, y->d, dst))) return rc;
    if ((rc = fetch(c, dst, out, 1))) return rc;
    if (x->dtype == HFH_F32) *out = static_cast<float>(*out);
    return HFH_OK;
}

extern "C" int hfh_vec_device_ptr(const hfh_vec *v, void **ptr) {
    HFH_CHECK(v && ptr, HFH_ERR_INVALID, "hfh_vec_device_ptr: null argument");
    *ptr = v->d;
    return HFH_OK;
}

extern "C" int hfh_dot_dev(const hfh_vec *x, const hfh_vec *y, double *dev_out) {
    int rc = same_shape(x, y, "hfh_dot_dev");
    if (rc) return rc;
    HFH_CHECK(dev_out, HFH_ERR_INVALID, "hfh_dot_dev: null output");
    hfh_ctx *c = x->ctx;
    if (x->n == 0) {
        // no partials: the finisher stores 0
        hipLaunchKernelGGL(k_fin_store, dim3(1), dim3(kFinBlock), 0, c->stream, (const double *)c->red, 0, dev_out);
        return launch_check("k_fin_store");
    }
    return dot_async(c, x->dtype, x->n, x->d, y->d, dev_out);
}

extern "C" int hfh_nrm2(const hfh_vec *x, double *out) {
    int rc = hfh_dot(x, x, out);
    if (rc) return rc;
    *out = x->dtype == HFH_F32 ? static_cast<double>(sqrtf(static_cast<float>(*out))) : sqrt(*out);
    return HFH_OK;
}

// Launches y = op(y, x) (x may be null for a unary op).
template <template <typename> class Op, bool BIN>
static int launch_update(hfh_vec *y, const hfh_vec *x, double a, const char *what) {
    // operands beyond the Infinity Cache stream with non-temporal accesses
    // from more workgroups
    const bool big = (BIN ? 2.0 : 1.0) * y->n * y->elem() > kMallBytes;
    const int nb = big ? red_blocks(y->n, kStreamBlocksBig) : ew_blocks(y->ctx, y->n);
    hipStream_t s = y->ctx->stream;
    const void *xd = BIN ? x->d : y->d;
    if (y->dtype == HFH_F64) {
        Op<double> op{};
        set_scalar(op, a);
        if (big)
            hipLaunchKernelGGL((k_update<double, Op<double>, BIN, true>), dim3(nb), dim3(kBlock), 0, s, y->n,
                               vec_count<double>(y->n, y->d, xd), (double *)y->d, (const double *)xd, op);
        else
            hipLaunchKernelGGL((k_update<double, Op<double>, BIN>), dim3(nb), dim3(kBlock), 0, s, y->n,
                               vec_count<double>(y->n, y->d, xd), (double *)y->d, (const double *)xd, op);
    } else {
        Op<float> op{};
        set_scalar(op, a);
        if (big)
            hipLaunchKernelGGL((k_update<float, Op<float>, BIN, true>), dim3(nb), dim3(kBlock), 0, s, y->n,
                               vec_count<float>(y->n, y->d, xd), (float *)y->d, (const float *)xd, op);
        else
            hipLaunchKernelGGL((k_update<float, Op<float>, BIN>), dim3(nb), dim3(kBlock), 0, s, y->n,
                               vec_count<float>(y->n, y->d, xd), (float *)y->d, (const float *)xd, op);
    }
    return launch_check(what);
}

extern "C" int hfh_axpy(hfh_vec *y, const hfh_vec *x, double a) {
    int rc = same_shape(y, x, "hfh_axpy");
    if (rc || y->n == 0) return rc;
    return launch_update<OpAxpy, true>(y, x, a, "hfh_axpy");
}

extern "C" int hfh_scaleadd(hfh_vec *y, double a, const hfh_vec *x) {
    int rc = same_shape(y, x, "hfh_scaleadd");
    if (rc || y->n == 0) return rc;
    return launch_update<OpScaleAdd, true>(y, x, a, "hfh_scaleadd");
}

extern "C" int hfh_scale(hfh_vec *y, double a) {
    HFH_CHECK(y, HFH_ERR_INVALID, "hfh_scale: null vector");
    if (y->n == 0) return HFH_OK;
    return launch_update<OpScale, false>(y, nullptr, a, "hfh_scale");
}

extern "C" int hfh_ewmult(hfh_vec *y, const hfh_vec *x) {
    int rc = same_shape(y, x, "hfh_ewmult");
    if (rc || y->n == 0) return rc;
    return launch_update<OpMult, true>(y, x, 0.0, "hfh_ewmult");
}
