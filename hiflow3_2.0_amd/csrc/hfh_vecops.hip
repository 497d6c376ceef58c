// hfh_vecops.hip -- the rest of lVector<T>'s interface on the gfx950 platform
// (lmp/lvector.h:39-336): Norm1, ArgMin / ArgMax / NormMax, add_value(s),
// the index set and its gather / scatter, partial_add_subvector, Rot and
// Rotm.  Same conventions as hfh_core.hip: deterministic partials per
// workgroup and a one-workgroup finisher, element-wise updates rounded
// exactly as the reference's CPU loops.
#include "hfh_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace hfh {
namespace {

constexpr int kAbsBlocks = 4096;   // workgroups (= partials) of Norm1 / ArgMin / ArgMax

inline int abs_blocks(int n) {
    long long nb = (static_cast<long long>(n) + kBlock - 1) / kBlock;
    return static_cast<int>(std::max<long long>(1, std::min<long long>(nb, kAbsBlocks)));
}

inline int ew_grid(const hfh_ctx *c, long long n) {
    long long nb = (n + kBlock - 1) / kBlock;
    const long long cap = static_cast<long long>(c->num_cu > 0 ? c->num_cu : 256) * 16;
    return static_cast<int>(std::max<long long>(1, std::min(nb, cap)));
}

// Sum of |x_i| in FP64 (Norm1, CPUsimple_lvector.cc:95-140); partials as k_dot.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_asum(int n, const T *__restrict__ x, double *part) {
    __shared__ double sh[kBlock / 64];
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    double acc[2] = {0, 0};
    long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x;
    for (; i + stride < n; i += 2 * stride) {
        const T a = x[i], b = x[i + stride];
        acc[0] += std::fabs(static_cast<double>(a));
        acc[1] += std::fabs(static_cast<double>(b));
    }
    if (i < n) acc[0] += std::fabs(static_cast<double>(x[i]));
    const double v[1] = {block_sum(acc[0] + acc[1], sh)};
    store_partials<1>(v, part, kRedBlocks, blockIdx.x);
}

// Candidate of ArgMax (MAX) / ArgMin over |x_i|: NaN entries are never
// selected, ties keep the lower index.  Index -1: no candidate.
template <bool MAX>
__device__ __forceinline__ bool better(double v, long long i, double bv, long long bi) {
    if (bi < 0) return true;
    if (MAX ? v > bv : v < bv) return true;
    return v == bv && i < bi;
}

// Per workgroup: the best (|x_i|, i) over its grid-stride elements, stored as
// value at part[b] and index at part[kAbsBlocks + b] (exact in FP64).
template <typename T, bool MAX>
__global__ __launch_bounds__(kBlock) void k_iabs(int n, const T *__restrict__ x, double *part) {
    __shared__ double sv[kBlock];
    __shared__ long long si[kBlock];
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    double bv = 0;
    long long bi = -1;
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < n; i += stride) {
        const double v = std::fabs(static_cast<double>(x[i]));
        if (v == v && better<MAX>(v, i, bv, bi)) {
            bv = v;
            bi = i;
        }
    }
    sv[threadIdx.x] = bv;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            const long long oi = si[threadIdx.x + h];
            if (oi >= 0 && better<MAX>(sv[threadIdx.x + h], oi, sv[threadIdx.x], si[threadIdx.x])) {
                sv[threadIdx.x] = sv[threadIdx.x + h];
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = sv[0];
        part[kAbsBlocks + blockIdx.x] = static_cast<double>(si[0]);
    }
}

// Finisher: the best of nb partials; then the reference's start value: the
// sequential loop begins with |x_0| and tests current > val (current < val
// for ArgMin), so a NaN x_0 is never replaced and the answer is 0, and an
// all-NaN vector answers 0 as well.  Result (index) at dst[0].
template <typename T, bool MAX>
__global__ __launch_bounds__(kFinBlock) void k_iabs_fin(const double *part, int nb, const T *__restrict__ x,
                                                        double *dst) {
    __shared__ double sv[kFinBlock];
    __shared__ long long si[kFinBlock];
    double bv = 0;
    long long bi = -1;
    for (int b = threadIdx.x; b < nb; b += kFinBlock) {
        const long long i = static_cast<long long>(part[kAbsBlocks + b]);
        if (i >= 0 && better<MAX>(part[b], i, bv, bi)) {
            bv = part[b];
            bi = i;
        }
    }
    sv[threadIdx.x] = bv;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int h = kFinBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            const long long oi = si[threadIdx.x + h];
            if (oi >= 0 && better<MAX>(sv[threadIdx.x + h], oi, sv[threadIdx.x], si[threadIdx.x])) {
                sv[threadIdx.x] = sv[threadIdx.x + h];
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double x0 = std::fabs(static_cast<double>(x[0]));
        dst[0] = (x0 != x0 || si[0] < 0) ? 0.0 : static_cast<double>(si[0]);
    }
}

// add_values (lvector_cpu.cc:123-135): the indices grouped on the host, each
// group's values in their original order; one thread per distinct index adds
// them one after the other, as the sequential loop does.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_add_grouped(int ng, T *__restrict__ y, const int *__restrict__ gidx,
                                                        const int *__restrict__ goff, const T *__restrict__ vals) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= ng) return;
    T t = y[gidx[g]];
    for (int j = goff[g]; j < goff[g + 1]; ++j) t = add_rn(t, vals[j]);
    y[gidx[g]] = t;
}

// partial_add_subvector (lvector_cpu.cc:532-548): y[i] += w * x[i], rounded separately.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_add_range(long long n, T *__restrict__ y, const T *__restrict__ x, T w) {
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * kBlock)
        y[i] = add_rn(y[i], mul_rn(w, x[i]));
}

// Rot (CPUsimple_lvector.cc:448-466): t = c a + s b; b = c b - s a; a = t.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_rot(int n, T *__restrict__ a, T *__restrict__ b, T c, T s) {
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * kBlock) {
        const T ai = a[i], bi = b[i];
        a[i] = add_rn(mul_rn(c, ai), mul_rn(s, bi));
        b[i] = add_rn(mul_rn(c, bi), -mul_rn(s, ai));
    }
}

// Rotm (CPUsimple_lvector.cc:540-600) with the reference's roles: the
// argument vector v is the first operand, this (u) the second;
// flag -1: v' = v h11 + u h12, u' = v h21 + u h22; flag 0: v' = v + u h12,
// u' = v h21 + u; flag 1: v' = v h11 + u, u' = u h22 - v.
template <typename T, int FLAG>
__global__ __launch_bounds__(kBlock) void k_rotm(int n, T *__restrict__ u, T *__restrict__ v, T h11, T h21, T h12,
                                                 T h22) {
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * kBlock) {
        const T t1 = v[i], t2 = u[i];
        if (FLAG == -1) {
            v[i] = add_rn(mul_rn(t1, h11), mul_rn(t2, h12));
            u[i] = add_rn(mul_rn(t1, h21), mul_rn(t2, h22));
        } else if (FLAG == 0) {
            v[i] = add_rn(t1, mul_rn(t2, h12));
            u[i] = add_rn(mul_rn(t1, h21), t2);
        } else {
            v[i] = add_rn(mul_rn(t1, h11), t2);
            u[i] = add_rn(mul_rn(t2, h22), -t1);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather_v(int n, T *__restrict__ y, const T *__restrict__ x,
                                                     const int *__restrict__ idx) {
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * kBlock)
        y[i] = x[idx[i]];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scatter_v(int n, T *__restrict__ y, const T *__restrict__ x,
                                                      const int *__restrict__ idx) {
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * kBlock)
        y[idx[i]] = x[i];
}

int check_vec(const hfh_vec *v, const char *what) {
    HFH_CHECK(v, HFH_ERR_INVALID, "%s: null vector", what);
    return HFH_OK;
}

// Device scratch freed when the call returns (after its stream work).
struct Scratch {
    std::vector<void *> p;
    hipStream_t s = nullptr;
    ~Scratch() {
        if (!p.empty() && s) (void)hipStreamSynchronize(s);
        for (void *q : p) (void)hipFree(q);
    }
    int get(void **out, size_t bytes) {
        int rc = dev_alloc(out, bytes);
        if (!rc) p.push_back(*out);
        return rc;
    }
};

// Positions of `index` whose write is the last one to its element: a scatter
// of only these gives the result of the reference's sequential SetValues loop
// (lvector_cpu.cc:137-151) when an index repeats.
std::vector<int> last_writes(const int *index, int count, int n) {
    std::vector<int> last(n, -1), keep;
    bool dup = false;
    for (int i = 0; i < count; ++i) {
        if (last[index[i]] >= 0) dup = true;
        last[index[i]] = i;
    }
    keep.reserve(count);
    for (int i = 0; i < count; ++i)
        if (!dup || last[index[i]] == i) keep.push_back(i);
    return keep;
}

int scatter_host(hfh_vec *v, const int *index, int count, const void *values, const char *what) {
    if (count == 0) return HFH_OK;
    const std::vector<int> keep = last_writes(index, count, v->n);
    const int m = static_cast<int>(keep.size());
    std::vector<int> idx(m);
    std::vector<char> val(v->elem() * m);
    for (int k = 0; k < m; ++k) {
        idx[k] = index[keep[k]];
        std::memcpy(&val[v->elem() * k], static_cast<const char *>(values) + v->elem() * keep[k], v->elem());
    }
    Scratch S;
    S.s = v->ctx->stream;
    void *dv = nullptr, *di = nullptr;
    int rc;
    if ((rc = S.get(&dv, v->elem() * m)) || (rc = S.get(&di, sizeof(int) * m))) return rc;
    HFH_HIP(hipMemcpyAsync(dv, val.data(), v->elem() * m, hipMemcpyHostToDevice, S.s));
    HFH_HIP(hipMemcpyAsync(di, idx.data(), sizeof(int) * m, hipMemcpyHostToDevice, S.s));
    const int nb = ew_grid(v->ctx, m);
    if (v->dtype == HFH_F64)
        hipLaunchKernelGGL(k_scatter_v<double>, dim3(nb), dim3(kBlock), 0, S.s, m, (double *)v->d, (const double *)dv, (const int *)di);
    else
        hipLaunchKernelGGL(k_scatter_v<float>, dim3(nb), dim3(kBlock), 0, S.s, m, (float *)v->d, (const float *)dv, (const int *)di);
    return launch_check(what);
}

}  // namespace
}  // namespace hfh

using namespace hfh;

extern "C" int hfh_norm1(const hfh_vec *x, double *out) {
    if (int rc = check_vec(x, "hfh_norm1")) return rc;
    HFH_CHECK(out, HFH_ERR_INVALID, "hfh_norm1: null output");
    HFH_CHECK(x->n > 0, HFH_ERR_INVALID, "hfh_norm1: empty vector");
    hfh_ctx *c = x->ctx;
    const int nb = abs_blocks(x->n);
    if (x->dtype == HFH_F64)
        hipLaunchKernelGGL(k_asum<double>, dim3(nb), dim3(kBlock), 0, c->stream, x->n, (const double *)x->d, c->red);
    else
        hipLaunchKernelGGL(k_asum<float>, dim3(nb), dim3(kBlock), 0, c->stream, x->n, (const float *)x->d, c->red);
    int rc = launch_check("k_asum");
    if (rc) return rc;
    double *dst = c->red + kRedBlocks * 8;
    if ((rc = partials_sum(c, nb, dst))) return rc;
    if ((rc = fetch(c, dst, out, 1))) return rc;
    if (x->dtype == HFH_F32) *out = static_cast<float>(*out);
    return HFH_OK;
}

template <bool MAX>
static int iabs(const hfh_vec *x, int *out, const char *what) {
    if (int rc = check_vec(x, what)) return rc;
    HFH_CHECK(out, HFH_ERR_INVALID, "%s: null output", what);
    HFH_CHECK(x->n > 0, HFH_ERR_INVALID, "%s: empty vector", what);
    hfh_ctx *c = x->ctx;
    const int nb = abs_blocks(x->n);
    static_assert(2 * kAbsBlocks <= kRedBlocks * 8, "index partials must fit the context's partial region");
    double *dst = c->red + kRedBlocks * 8;
    if (x->dtype == HFH_F64) {
        hipLaunchKernelGGL((k_iabs<double, MAX>), dim3(nb), dim3(kBlock), 0, c->stream, x->n, (const double *)x->d, c->red);
        hipLaunchKernelGGL((k_iabs_fin<double, MAX>), dim3(1), dim3(kFinBlock), 0, c->stream, (const double *)c->red, nb,
                           (const double *)x->d, dst);
    } else {
        hipLaunchKernelGGL((k_iabs<float, MAX>), dim3(nb), dim3(kBlock), 0, c->stream, x->n, (const float *)x->d, c->red);
        hipLaunchKernelGGL((k_iabs_fin<float, MAX>), dim3(1), dim3(kFinBlock), 0, c->stream, (const double *)c->red, nb,
                           (const float *)x->d, dst);
    }
    int rc = launch_check(what);
    if (rc) return rc;
    double r = 0;
    if ((rc = fetch(c, dst, &r, 1))) return rc;
    *out = static_cast<int>(r);
    return HFH_OK;
}

extern "C" int hfh_iamax(const hfh_vec *x, int *index) { return iabs<true>(x, index, "hfh_iamax"); }
extern "C" int hfh_iamin(const hfh_vec *x, int *index) { return iabs<false>(x, index, "hfh_iamin"); }

extern "C" int hfh_normmax(const hfh_vec *x, double *out) {
    HFH_CHECK(out, HFH_ERR_INVALID, "hfh_normmax: null output");
    int i = 0;
    int rc = hfh_iamax(x, &i);
    if (rc) return rc;
    if (x->dtype == HFH_F64) {
        double v = 0;
        if ((rc = hfh_vec_download(x, &v, i, 1))) return rc;
        *out = std::fabs(v);
    } else {
        float v = 0;
        if ((rc = hfh_vec_download(x, &v, i, 1))) return rc;
        *out = std::fabs(v);
    }
    return HFH_OK;
}

extern "C" int hfh_vec_add_values(hfh_vec *v, const int *index, int count, const void *values) {
    if (int rc = check_vec(v, "hfh_vec_add_values")) return rc;
    HFH_CHECK(count >= 0 && (count == 0 || (index && values)), HFH_ERR_INVALID, "hfh_vec_add_values: bad argument");
    if (count == 0) return HFH_OK;
    for (int i = 0; i < count; ++i)
        HFH_CHECK(index[i] >= 0 && index[i] < v->n, HFH_ERR_INVALID, "hfh_vec_add_values: index[%d]=%d outside [0,%d)",
                  i, index[i], v->n);
    // group the entries by index, keeping their order inside a group
    std::vector<int> ord(count);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return index[a] < index[b]; });
    std::vector<int> gidx, goff;
    std::vector<char> vals(v->elem() * count);
    for (int k = 0; k < count; ++k) {
        const int i = ord[k];
        if (k == 0 || index[i] != index[ord[k - 1]]) {
            gidx.push_back(index[i]);
            goff.push_back(k);
        }
        std::memcpy(&vals[v->elem() * k], static_cast<const char *>(values) + v->elem() * i, v->elem());
    }
    goff.push_back(count);
    const int ng = static_cast<int>(gidx.size());
    Scratch S;
    S.s = v->ctx->stream;
    void *dgi = nullptr, *dgo = nullptr, *dv = nullptr;
    int rc;
    if ((rc = S.get(&dgi, sizeof(int) * ng)) || (rc = S.get(&dgo, sizeof(int) * (ng + 1))) ||
        (rc = S.get(&dv, v->elem() * count)))
        return rc;
    HFH_HIP(hipMemcpyAsync(dgi, gidx.data(), sizeof(int) * ng, hipMemcpyHostToDevice, S.s));
    HFH_HIP(hipMemcpyAsync(dgo, goff.data(), sizeof(int) * (ng + 1), hipMemcpyHostToDevice, S.s));
    HFH_HIP(hipMemcpyAsync(dv, vals.data(), v->elem() * count, hipMemcpyHostToDevice, S.s));
    const int nb = (ng + kBlock - 1) / kBlock;
    if (v->dtype == HFH_F64)
        hipLaunchKernelGGL(k_add_grouped<double>, dim3(nb), dim3(kBlock), 0, S.s, ng, (double *)v->d, (const int *)dgi,
                           (const int *)dgo, (const double *)dv);
    else
        hipLaunchKernelGGL(k_add_grouped<float>, dim3(nb), dim3(kBlock), 0, S.s, ng, (float *)v->d, (const int *)dgi,
                           (const int *)dgo, (const float *)dv);
    return launch_check("k_add_grouped");
}

extern "C" int hfh_vec_add_range(hfh_vec *dst, int doff, const hfh_vec *src, int soff, int count, double weight) {
    HFH_CHECK(dst && src, HFH_ERR_INVALID, "hfh_vec_add_range: null vector");
    HFH_CHECK(dst->dtype == src->dtype && dst->ctx == src->ctx, HFH_ERR_INVALID, "hfh_vec_add_range: dtype mismatch");
    HFH_CHECK(doff >= 0 && soff >= 0 && count >= 0 && static_cast<long long>(doff) + count <= dst->n &&
                  static_cast<long long>(soff) + count <= src->n,
              HFH_ERR_INVALID, "hfh_vec_add_range: range out of bounds");
    HFH_CHECK(dst->d != src->d || doff == soff || doff + count <= soff || soff + count <= doff, HFH_ERR_INVALID,
              "hfh_vec_add_range: overlapping ranges of one vector");
    if (count == 0) return HFH_OK;
    const int nb = ew_grid(dst->ctx, count);
    hipStream_t s = dst->ctx->stream;
    if (dst->dtype == HFH_F64)
        hipLaunchKernelGGL(k_add_range<double>, dim3(nb), dim3(kBlock), 0, s, count, (double *)dst->d + doff,
                           (const double *)src->d + soff, weight);
    else
        hipLaunchKernelGGL(k_add_range<float>, dim3(nb), dim3(kBlock), 0, s, count, (float *)dst->d + doff,
                           (const float *)src->d + soff, static_cast<float>(weight));
    return launch_check("k_add_range");
}

extern "C" int hfh_vec_set_indexset(hfh_vec *v, const int *index, int count) {
    if (int rc = check_vec(v, "hfh_vec_set_indexset")) return rc;
    HFH_CHECK(count >= 0 && (count == 0 || index), HFH_ERR_INVALID, "hfh_vec_set_indexset: bad argument");
    for (int i = 0; i < count; ++i)
        HFH_CHECK(index[i] >= 0 && index[i] < v->n, HFH_ERR_INVALID,
                  "hfh_vec_set_indexset: index[%d]=%d outside [0,%d)", i, index[i], v->n);
    if (v->iset) {
        HFH_HIP(hipStreamSynchronize(v->ctx->stream));
        (void)hipFree(v->iset);
        v->iset = nullptr;
    }
    v->h_iset.assign(index, index + count);
    if (count == 0) return HFH_OK;
    int rc = dev_alloc(reinterpret_cast<void **>(&v->iset), sizeof(int) * count);
    if (rc) {
        v->h_iset.clear();
        return rc;
    }
    HFH_HIP(hipMemcpyAsync(v->iset, index, sizeof(int) * count, hipMemcpyHostToDevice, v->ctx->stream));
    HFH_HIP(hipStreamSynchronize(v->ctx->stream));
    return HFH_OK;
}

extern "C" int hfh_vec_get_indexset(const hfh_vec *v, int *index, int *count) {
    if (int rc = check_vec(v, "hfh_vec_get_indexset")) return rc;
    HFH_CHECK(count, HFH_ERR_INVALID, "hfh_vec_get_indexset: null count");
    *count = static_cast<int>(v->h_iset.size());
    if (index) std::copy(v->h_iset.begin(), v->h_iset.end(), index);
    return HFH_OK;
}

extern "C" int hfh_vec_get_indexed_values(const hfh_vec *v, void *values) {
    if (int rc = check_vec(v, "hfh_vec_get_indexed_values")) return rc;
    const int m = static_cast<int>(v->h_iset.size());
    if (m == 0 || v->n == 0) return HFH_OK;
    HFH_CHECK(values, HFH_ERR_INVALID, "hfh_vec_get_indexed_values: null output");
    Scratch S;
    S.s = v->ctx->stream;
    void *dv = nullptr;
    int rc = S.get(&dv, v->elem() * m);
    if (rc) return rc;
    const int nb = ew_grid(v->ctx, m);
    if (v->dtype == HFH_F64)
        hipLaunchKernelGGL(k_gather_v<double>, dim3(nb), dim3(kBlock), 0, S.s, m, (double *)dv, (const double *)v->d, v->iset);
    else
        hipLaunchKernelGGL(k_gather_v<float>, dim3(nb), dim3(kBlock), 0, S.s, m, (float *)dv, (const float *)v->d, v->iset);
    if ((rc = launch_check("k_gather_v"))) return rc;
    HFH_HIP(hipMemcpyAsync(values, dv, v->elem() * m, hipMemcpyDeviceToHost, S.s));
    HFH_HIP(hipStreamSynchronize(S.s));
    return HFH_OK;
}

extern "C" int hfh_vec_set_indexed_values(hfh_vec *v, const void *values) {
    if (int rc = check_vec(v, "hfh_vec_set_indexed_values")) return rc;
    const int m = static_cast<int>(v->h_iset.size());
    if (m == 0 || v->n == 0 || !values) return HFH_OK;
    return scatter_host(v, v->h_iset.data(), m, values, "hfh_vec_set_indexed_values");
}

extern "C" int hfh_vec_set_values(hfh_vec *v, const int *index, int count, const void *values) {
    if (int rc = check_vec(v, "hfh_vec_set_values")) return rc;
    HFH_CHECK(count >= 0 && (count == 0 || (index && values)), HFH_ERR_INVALID, "hfh_vec_set_values: bad argument");
    for (int i = 0; i < count; ++i)
        HFH_CHECK(index[i] >= 0 && index[i] < v->n, HFH_ERR_INVALID, "hfh_vec_set_values: index[%d]=%d outside [0,%d)",
                  i, index[i], v->n);
    return scatter_host(v, index, count, values, "hfh_vec_set_values");
}

extern "C" int hfh_rot(hfh_vec *x, hfh_vec *y, double c, double s) {
    HFH_CHECK(x && y, HFH_ERR_INVALID, "hfh_rot: null vector");
    HFH_CHECK(x->n == y->n && x->dtype == y->dtype && x->ctx == y->ctx, HFH_ERR_INVALID,
              "hfh_rot: vectors differ in size/dtype/context");
    HFH_CHECK(x->d != y->d, HFH_ERR_INVALID, "hfh_rot: x and y must not alias");
    if (x->n == 0) return HFH_OK;
    const int nb = ew_grid(x->ctx, x->n);
    hipStream_t st = x->ctx->stream;
    if (x->dtype == HFH_F64)
        hipLaunchKernelGGL(k_rot<double>, dim3(nb), dim3(kBlock), 0, st, x->n, (double *)x->d, (double *)y->d, c, s);
    else
        hipLaunchKernelGGL(k_rot<float>, dim3(nb), dim3(kBlock), 0, st, x->n, (float *)x->d, (float *)y->d,
                           static_cast<float>(c), static_cast<float>(s));
    return launch_check("k_rot");
}

template <typename T>
static void launch_rotm(hfh_vec *u, hfh_vec *v, const double *p, int flag) {
    const int nb = ew_grid(u->ctx, u->n);
    hipStream_t st = u->ctx->stream;
    const T h11 = static_cast<T>(p[1]), h21 = static_cast<T>(p[2]), h12 = static_cast<T>(p[3]),
            h22 = static_cast<T>(p[4]);
    if (flag == -1)
        hipLaunchKernelGGL((k_rotm<T, -1>), dim3(nb), dim3(kBlock), 0, st, u->n, (T *)u->d, (T *)v->d, h11, h21, h12, h22);
    else if (flag == 0)
        hipLaunchKernelGGL((k_rotm<T, 0>), dim3(nb), dim3(kBlock), 0, st, u->n, (T *)u->d, (T *)v->d, h11, h21, h12, h22);
    else
        hipLaunchKernelGGL((k_rotm<T, 1>), dim3(nb), dim3(kBlock), 0, st, u->n, (T *)u->d, (T *)v->d, h11, h21, h12, h22);
}

extern "C" int hfh_rotm(hfh_vec *u, hfh_vec *v, const double *param) {
    HFH_CHECK(u && v && param, HFH_ERR_INVALID, "hfh_rotm: null argument");
    HFH_CHECK(u->n == v->n && u->dtype == v->dtype && u->ctx == v->ctx, HFH_ERR_INVALID,
              "hfh_rotm: vectors differ in size/dtype/context");
    HFH_CHECK(u->d != v->d, HFH_ERR_INVALID, "hfh_rotm: vectors must not alias");
    // the flag as the reference compares it, in the vector's type
    const double flag = u->dtype == HFH_F64 ? param[0] : static_cast<double>(static_cast<float>(param[0]));
    if (u->n == 0 || !(flag == -1.0 || flag == 0.0 || flag == 1.0)) return HFH_OK;   // -2 and others: identity
    if (u->dtype == HFH_F64) launch_rotm<double>(u, v, param, static_cast<int>(flag));
    else launch_rotm<float>(u, v, param, static_cast<int>(flag));
    return launch_check("k_rotm");
}
