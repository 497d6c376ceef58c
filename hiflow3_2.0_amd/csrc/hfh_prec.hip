// hfh_prec.hip -- preconditioner applies on gfx950.
//
// Jacobi  : out = in .* invD (lpreconditioner.cc:151-166), one HBM stream.
// MC-SGS  : on the colour-permuted matrix every colour is a contiguous row
//           range and rows of one colour do not couple, so each colour's
//           forward / backward block step (lpreconditioner_mc_gs.cc:91-113,
//           lpreconditioner_mc_sgs.cc:71-90) is one data-parallel launch over
//           that colour's rows; the rows read only the off-colour part of A.
//           Block products are summed per colour block and added in the
//           reference's block order (ascending forward, descending backward).
// SGS / ILU(0): sparse triangular sweeps (lmatrix_csr_cpu.cc:542-616 and
//           2108-2171) scheduled by levels: rows of one level are
//           independent, one launch per level, with each row's arithmetic in
//           the reference's order.
#include "hfh_internal.h"

#include <algorithm>
#include <type_traits>

namespace hfh {

constexpr int kMaxColors = 64;

constexpr int kMergeBlock = 1024;   // threads of a merged-level workgroup; levels up to this size merge

struct ColorOffsets {
    int off[kMaxColors + 1];
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_jacobi(int n, const T *__restrict__ in, const T *__restrict__ invd,
                                                   T *__restrict__ out, const int *__restrict__ done) {
    if (done && *done) return;
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        out[i] = in[i] * invd[i];
}

// forward step of colour 0 (no lower colours): y = (-in) * iDv
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mc_fwd0(int r0, int r1, const T *__restrict__ in,
                                                    const T *__restrict__ idv, T *__restrict__ out,
                                                    const int *__restrict__ done) {
    if (done && *done) return;
    const int i = r0 + blockIdx.x * kBlock + threadIdx.x;
    if (i >= r1) return;
    out[i] = (in[i] * static_cast<T>(-1)) * idv[i];
}

// forward step of colour c >= 1.  When `last` the colour's diagonal and
// backward steps (no upper colours) are applied too: z = ((y * Dv) * iDv).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mc_fwd(int c, ColorOffsets co, bool last,
                                                   const int *__restrict__ row, const int *__restrict__ col,
                                                   const T *__restrict__ val, const T *__restrict__ in,
                                                   const T *__restrict__ idv, const T *__restrict__ dv,
                                                   T *__restrict__ out, const int *__restrict__ done) {
    if (done && *done) return;
    const int i = co.off[c] + blockIdx.x * kBlock + threadIdx.x;
    if (i >= co.off[c + 1]) return;
    T chunk = in[i] * static_cast<T>(-1);
    int blk = 0;
    int blk_end = co.off[1];
    T res = 0;
    const int lo = co.off[c];
    for (int j = row[i], je = row[i + 1]; j < je; ++j) {
        const int cj = col[j];
        if (cj >= lo) break;
        while (cj >= blk_end) {  // entries are column sorted: flush finished blocks
            chunk += res;
            res = 0;
            ++blk;
            blk_end = co.off[blk + 1];
        }
        res += val[j] * out[cj];
    }
    chunk += res;
    chunk *= idv[i];
    if (last) chunk = (chunk * dv[i]) * idv[i];
    out[i] = chunk;
}

// backward step of colour c < ncolors-1, with the diagonal step first.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mc_bwd(int c, int ncolors, ColorOffsets co,
                                                   const int *__restrict__ row, const int *__restrict__ col,
                                                   const T *__restrict__ val, const T *__restrict__ idv,
                                                   const T *__restrict__ dv, T *__restrict__ out,
                                                   const int *__restrict__ done) {
    if (done && *done) return;
    const int i = co.off[c] + blockIdx.x * kBlock + threadIdx.x;
    if (i >= co.off[c + 1]) return;
    T chunk = out[i] * dv[i];
    const int rs = row[i];
    int e = row[i + 1];
    for (int b = ncolors - 1; b > c; --b) {
        const int lo = co.off[b];
        int s = e;
        while (s > rs && col[s - 1] >= lo) --s;
        T res = 0;
        for (int j = s; j < e; ++j) res += val[j] * out[col[j]];
        chunk += res;
        e = s;
    }
    out[i] = chunk * idv[i];
}

// The same colour steps on the SELL-64 copies of the sweeps: one wavefront
// per slice of 64 rows of colour c, lane = row, so the value/column streams
// and the gathers of neighbouring rows are contiguous.  Each row's entries
// and its block-by-block summation are those of k_mc_fwd / k_mc_bwd.
constexpr int kSweepU = 8;   // column slots in flight (two code words when coded)

// Columns of the kSweepU slots k.. of row i: stored (scol), or row + the
// offset coded in the slot's byte of its code word (codes: kSweepU / 4 words
// per step; the column clamped to [0, n): padding slots and rows past the
// colour read a valid entry, and their products are not summed).  live[u]:
// slot k + u is one of the row's entries (stored: k + u < len; coded: its
// code is not kPadCode, so the row's length is not loaded).  Slots past the
// slice's width L are not loaded (a wave-uniform test: L is the slice's), so
// a slice of 3 slots issues 3 loads, not kSweepU.
// The loads (sweep_raw: stored columns, or kSweepU / 4 code words) and their
// decoding (sweep_decode) are separate so that a wave can issue the loads of
// its next slice before it needs them.
template <bool CODED>
__device__ __forceinline__ void sweep_raw(unsigned (&raw)[kSweepU], int k, int L, int base, int wbase, int lane, int i,
                                          int n, const int *__restrict__ scol, const unsigned *__restrict__ code) {
    if (!CODED) {
#pragma unroll
        for (int u = 0; u < kSweepU; ++u)
            raw[u] = k + u < L ? __builtin_nontemporal_load(&scol[base + (k + u) * 64 + lane]) : min(i, n - 1);
        return;
    }
    const int W = (L + 3) >> 2;   // code words of the slice (its width rounded up to 4 slots)
#pragma unroll
    for (int q = 0; q < kSweepU / 4; ++q)
        raw[q] = (k >> 2) + q < W ? __builtin_nontemporal_load(&code[wbase + ((k >> 2) + q) * 64 + lane]) : ~0u;
}

// v, taken as unknown until this point: an empty asm that reads and writes
// it (and clobbers memory, so the loads before it stay before it) moves the
// wait for the load that produced v here, after those loads have issued.
template <typename T>
__device__ __forceinline__ T settled(T v) {
    asm volatile("" : "+v"(v) : : "memory");
    return v;
}

// sweep_raw without branches: every load reads a slot of the slice (the
// index clamped to its last slot, L > 0) and the slots past its width are
// replaced afterwards, so a step's loads issue back to back.
template <bool CODED>
__device__ __forceinline__ void sweep_raw_flat(unsigned (&raw)[kSweepU], int k, int L, int base, int wbase, int lane,
                                               int i, int n, const int *__restrict__ scol,
                                               const unsigned *__restrict__ code) {
    if (!CODED) {
#pragma unroll
        for (int u = 0; u < kSweepU; ++u) {
            const unsigned t = __builtin_nontemporal_load(&scol[base + min(k + u, L - 1) * 64 + lane]);
            raw[u] = k + u < L ? t : static_cast<unsigned>(min(i, n - 1));
        }
        return;
    }
    const int W = (L + 3) >> 2;
#pragma unroll
    for (int q = 0; q < kSweepU / 4; ++q) {
        const unsigned t = __builtin_nontemporal_load(&code[wbase + min((k >> 2) + q, W - 1) * 64 + lane]);
        raw[q] = (k >> 2) + q < W ? t : ~0u;
    }
}

template <bool CODED>
__device__ __forceinline__ void sweep_decode(const unsigned (&raw)[kSweepU], int (&cj)[kSweepU], bool (&live)[kSweepU],
                                             int k, int len, int i, int n, const int *dict) {
    if (!CODED) {
#pragma unroll
        for (int u = 0; u < kSweepU; ++u) {
            cj[u] = static_cast<int>(raw[u]);
            live[u] = k + u < len;
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < kSweepU / 4; ++q)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const unsigned cd = (raw[q] >> (8 * t)) & 255u;
            cj[4 * q + t] = min(i + dict[cd], n - 1);
            live[4 * q + t] = cd != kPadCode;
        }
}

template <bool CODED>
__device__ __forceinline__ void sweep_cols(int (&cj)[kSweepU], bool (&live)[kSweepU], int k, int L, int len, int base,
                                           int wbase, int lane, int i, int n, const int *__restrict__ scol,
                                           const unsigned *__restrict__ code, const int *dict) {
    unsigned raw[kSweepU];
    sweep_raw<CODED>(raw, k, L, base, wbase, lane, i, n, scol, code);
    sweep_decode<CODED>(raw, cj, live, k, len, i, n, dict);
}

// Workgroups per CU of a multicolour SGS colour step (0: one slice per
// wave).  FP32 steps are light: 8 workgroups per CU whose waves take several
// slices each (and store 8 r.z partials per CU instead of one per 4 slices)
// beat one slice per wave; FP64 keeps one slice per wave (FV 192^3: the FP32
// defect correction 96 -> 85 ms with the FP32 SpMV's like grid; FP64 CG +
// MC-SGS 150 ms either way, 158 ms with FP64 steps capped too).
template <typename T> constexpr int kSweepGridPerCu = 0;
template <> constexpr int kSweepGridPerCu<float> = 8;

// A slice's header, its first step's loads and its rows' own inputs.  On a
// capped grid (kSweepGridPerCu > 0) a wave walks several slices and issues
// these one slice ahead, after the current slice's gathers, so their latency
// overlaps the gathers.  With one slice per wave (FP64) the colour steps keep
// their loads in place: the look-ahead loop costs those kernels registers
// and occupancy (70 -> 92 VGPRs) for no overlap.
template <typename T>
struct SweepAhead {
    int base = 0, L = 0, wbase = 0, len = 0;
    unsigned raw[kSweepU];
    T v[kSweepU];
    T a = 0, b = 0;   // forward: r_i; backward: z_i and d_i
};

template <typename T, bool CODED>
__device__ __forceinline__ void sweep_ahead(SweepAhead<T> &na, int s, int i, bool valid, int lane, int n,
                                            const int *__restrict__ sptr, const int *__restrict__ slen,
                                            const int *__restrict__ scol, const unsigned *__restrict__ code,
                                            const int *__restrict__ wptr, const T *__restrict__ sval) {
    na.base = sptr[s];
    na.L = (sptr[s + 1] - na.base) >> 6;
    na.wbase = CODED ? wptr[s] : 0;
    na.len = !CODED && valid ? slen[i] : 0;   // coded: padding slots carry kPadCode
    sweep_raw<CODED>(na.raw, 0, na.L, na.base, na.wbase, lane, i, n, scol, code);
#pragma unroll
    for (int u = 0; u < kSweepU; ++u)   // slots past the slice width: not loaded (wave-uniform)
        na.v[u] = u < na.L ? __builtin_nontemporal_load(&sval[na.base + u * 64 + lane]) : static_cast<T>(0);
}

// RZ: the step also stores its workgroup's partial of in.out over its rows
// (r.z of a CG step, for the colour whose z this step finalises) in rzp.
template <typename T, bool CODED, bool RZ>
__global__ __launch_bounds__(kBlock) void k_mcs_fwd(int c, ColorOffsets co, bool last, int slice0, int nsl, int n,
                                                    const int *__restrict__ sptr, const int *__restrict__ slen,
                                                    const int *__restrict__ scol, const unsigned *__restrict__ code,
                                                    const int *__restrict__ wptr, const int *__restrict__ gdict,
                                                    int ndict, const T *__restrict__ sval, const T *__restrict__ in,
                                                    const T *__restrict__ idv, const T *__restrict__ dv, T *out,
                                                    const int *__restrict__ done, double *__restrict__ rzp) {
    if (done && *done) return;
    __shared__ int dict[CODED ? 256 : 1];
    if (CODED) {   // kPadCode: offset 0
        for (int t = threadIdx.x; t < 256; t += kBlock) dict[t] = t < ndict ? gdict[t] : 0;
        __syncthreads();
    }
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    double rz = 0;
    if constexpr (kSweepGridPerCu<T> == 0) {   // one slice per wave: loads in place
        for (int w = blockIdx.x * (kBlock / 64) + wave; w < nsl; w += gridDim.x * (kBlock / 64)) {
            const int s = slice0 + w;
            const int i = co.off[c] + w * 64 + lane;
            const bool valid = i < co.off[c + 1];
            const int base = sptr[s], L = (sptr[s + 1] - base) >> 6;
            const int wbase = CODED ? wptr[s] : 0;
            const int len = !CODED && valid ? slen[i] : 0;   // coded: padding slots carry kPadCode
            // r_i loaded unmasked (rows past the colour read row n-1; their
            // results are neither stored nor summed) and -r_i formed once the
            // first step's loads are issued, so no wait for r_i comes first
            const T ri = in[min(i, n - 1)];
            T chunk = 0;
            int blk = 0, blk_end = co.off[1];
            T res = 0;
            // one step of kSweepU slots; the first is peeled off the loop (a
            // slice of the stencil systems is one step), so its waits are not
            // merged with those of a loop back edge
            auto step = [&](int k) {
                int cj[kSweepU];
                bool live[kSweepU];
                unsigned raw[kSweepU];
                T v[kSweepU], xv[kSweepU];
                // straight-line loads: slots past the slice width re-read its
                // last slot (not live, not summed)
                sweep_raw_flat<CODED>(raw, k, L, base, wbase, lane, i, n, scol, code);
#pragma unroll
                for (int u = 0; u < kSweepU; ++u)
                    v[u] = __builtin_nontemporal_load(&sval[base + min(k + u, L - 1) * 64 + lane]);
                sweep_decode<CODED>(raw, cj, live, k, len, i, n, dict);
#pragma unroll
                for (int u = 0; u < kSweepU; ++u) xv[u] = out[cj[u]];
                if (k == 0) chunk = settled(ri) * static_cast<T>(-1);
#pragma unroll
                for (int u = 0; u < kSweepU; ++u) {
                    if (!live[u]) continue;
                    while (cj[u] >= blk_end) {   // entries are column sorted: flush finished blocks
                        chunk += res;
                        res = 0;
                        ++blk;
                        blk_end = co.off[blk + 1];
                    }
                    res += v[u] * xv[u];
                }
            };
            if (L > 0) {
                step(0);
                for (int k = kSweepU; k < L; k += kSweepU) step(k);
            } else {
                chunk = ri * static_cast<T>(-1);   // a slice without entries runs no step
            }
            if (valid) {
                chunk += res;
                chunk *= idv[i];
                if (last) chunk = (chunk * dv[i]) * idv[i];
                out[i] = chunk;
                if (RZ) rz += static_cast<double>(ri) * static_cast<double>(chunk);
            }
        }
    } else {
        const int stride = gridDim.x * (kBlock / 64);
        SweepAhead<T> na;
        auto ahead = [&](int wn) {
            const int in_ = co.off[c] + wn * 64 + lane;
            const bool vn = in_ < co.off[c + 1];
            sweep_ahead<T, CODED>(na, slice0 + wn, in_, vn, lane, n, sptr, slen, scol, code, wptr, sval);
            na.a = vn ? in[in_] : static_cast<T>(0);
        };
        int w = blockIdx.x * (kBlock / 64) + wave;
        if (w < nsl) ahead(w);
        for (; w < nsl; w += stride) {
            const int i = co.off[c] + w * 64 + lane;
            const bool valid = i < co.off[c + 1];
            const int base = na.base, L = na.L, wbase = na.wbase, len = na.len;
            const T ri = na.a;
            T chunk = ri * static_cast<T>(-1);
            int blk = 0, blk_end = co.off[1];
            T res = 0;
            for (int k = 0; k < L; k += kSweepU) {
                int cj[kSweepU];
                bool live[kSweepU];
                T v[kSweepU], xv[kSweepU];
                if (k == 0) {
                    sweep_decode<CODED>(na.raw, cj, live, 0, len, i, n, dict);
#pragma unroll
                    for (int u = 0; u < kSweepU; ++u) v[u] = na.v[u];
                } else {
                    sweep_cols<CODED>(cj, live, k, L, len, base, wbase, lane, i, n, scol, code, dict);
#pragma unroll
                    for (int u = 0; u < kSweepU; ++u)   // slots past the slice width: not loaded (wave-uniform)
                        v[u] = k + u < L ? __builtin_nontemporal_load(&sval[base + (k + u) * 64 + lane])
                                         : static_cast<T>(0);
                }
#pragma unroll
                for (int u = 0; u < kSweepU; ++u) xv[u] = k + u < L ? out[cj[u]] : static_cast<T>(0);
                if (k + kSweepU >= L && w + stride < nsl) ahead(w + stride);   // after this step's gathers
#pragma unroll
                for (int u = 0; u < kSweepU; ++u) {
                    if (!live[u]) continue;
                    while (cj[u] >= blk_end) {   // entries are column sorted: flush finished blocks
                        chunk += res;
                        res = 0;
                        ++blk;
                        blk_end = co.off[blk + 1];
                    }
                    res += v[u] * xv[u];
                }
            }
            if (L == 0 && w + stride < nsl) ahead(w + stride);   // an empty slice ran no step
            if (valid) {
                chunk += res;
                chunk *= idv[i];
                if (last) chunk = (chunk * dv[i]) * idv[i];
                out[i] = chunk;
                if (RZ) rz += static_cast<double>(ri) * static_cast<double>(chunk);
            }
        }
    }
    if (RZ) {
        __shared__ double sh[kBlock / 64];
        rz = block_sum(rz, sh);
        if (threadIdx.x == 0) rzp[blockIdx.x] = rz;
    }
}

// Backward step of colour c: entries stored block by block from the highest
// colour down, ascending columns within a block.
// RZ as in k_mcs_fwd, with the step's input r in rin.
template <typename T, bool CODED, bool RZ>
__global__ __launch_bounds__(kBlock) void k_mcs_bwd(int c, int ncolors, ColorOffsets co, int slice0, int nsl, int n,
                                                    const int *__restrict__ sptr, const int *__restrict__ slen,
                                                    const int *__restrict__ scol, const unsigned *__restrict__ code,
                                                    const int *__restrict__ wptr, const int *__restrict__ gdict,
                                                    int ndict, const T *__restrict__ sval, const T *__restrict__ idv,
                                                    const T *__restrict__ dv, T *out, const int *__restrict__ done,
                                                    const T *__restrict__ rin, double *__restrict__ rzp) {
    if (done && *done) return;
    __shared__ int dict[CODED ? 256 : 1];
    if (CODED) {   // kPadCode: offset 0
        for (int t = threadIdx.x; t < 256; t += kBlock) dict[t] = t < ndict ? gdict[t] : 0;
        __syncthreads();
    }
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    double rz = 0;
    if constexpr (kSweepGridPerCu<T> == 0) {   // one slice per wave: loads in place
        for (int w = blockIdx.x * (kBlock / 64) + wave; w < nsl; w += gridDim.x * (kBlock / 64)) {
            const int s = slice0 + w;
            const int i = co.off[c] + w * 64 + lane;
            const bool valid = i < co.off[c + 1];
            const int base = sptr[s], L = (sptr[s + 1] - base) >> 6;
            const int wbase = CODED ? wptr[s] : 0;
            const int len = !CODED && valid ? slen[i] : 0;   // coded: padding slots carry kPadCode
            // z_i and d_i loaded unmasked (rows past the colour read row n-1;
            // their results are not stored) and z_i d_i formed once the first
            // step's loads are issued
            const T zi0 = out[min(i, n - 1)], di = dv[min(i, n - 1)];
            T chunk = 0;
            int blk = ncolors - 1, blk_lo = co.off[ncolors - 1];
            T res = 0;
            // the first step peeled off the loop, as in k_mcs_fwd
            auto step = [&](int k) {
                int cj[kSweepU];
                bool live[kSweepU];
                unsigned raw[kSweepU];
                T v[kSweepU], xv[kSweepU];
                // straight-line loads: slots past the slice width re-read its
                // last slot (not live, not summed)
                sweep_raw_flat<CODED>(raw, k, L, base, wbase, lane, i, n, scol, code);
#pragma unroll
                for (int u = 0; u < kSweepU; ++u)
                    v[u] = __builtin_nontemporal_load(&sval[base + min(k + u, L - 1) * 64 + lane]);
                sweep_decode<CODED>(raw, cj, live, k, len, i, n, dict);
#pragma unroll
                for (int u = 0; u < kSweepU; ++u) xv[u] = out[cj[u]];
                if (k == 0) chunk = settled(zi0) * settled(di);
#pragma unroll
                for (int u = 0; u < kSweepU; ++u) {
                    if (!live[u]) continue;
                    while (cj[u] < blk_lo) {   // next lower colour block
                        chunk += res;
                        res = 0;
                        --blk;
                        blk_lo = co.off[blk];
                    }
                    res += v[u] * xv[u];
                }
            };
            if (L > 0) {
                step(0);
                for (int k = kSweepU; k < L; k += kSweepU) step(k);
            } else {
                chunk = zi0 * di;   // a slice without entries runs no step
            }
            if (valid) {
                chunk += res;
                const T zi = chunk * idv[i];
                out[i] = zi;
                if (RZ) rz += static_cast<double>(rin[i]) * static_cast<double>(zi);
            }
        }
    } else {
        const int stride = gridDim.x * (kBlock / 64);
        SweepAhead<T> na;
        auto ahead = [&](int wn) {   // z_i and d_i of the next slice's rows: only this wave writes them
            const int in_ = co.off[c] + wn * 64 + lane;
            const bool vn = in_ < co.off[c + 1];
            sweep_ahead<T, CODED>(na, slice0 + wn, in_, vn, lane, n, sptr, slen, scol, code, wptr, sval);
            na.a = vn ? out[in_] : static_cast<T>(0);
            na.b = vn ? dv[in_] : static_cast<T>(0);
        };
        int w = blockIdx.x * (kBlock / 64) + wave;
        if (w < nsl) ahead(w);
        for (; w < nsl; w += stride) {
            const int i = co.off[c] + w * 64 + lane;
            const bool valid = i < co.off[c + 1];
            const int base = na.base, L = na.L, wbase = na.wbase, len = na.len;
            T chunk = valid ? na.a * na.b : static_cast<T>(0);
            int blk = ncolors - 1, blk_lo = co.off[ncolors - 1];
            T res = 0;
            for (int k = 0; k < L; k += kSweepU) {
                int cj[kSweepU];
                bool live[kSweepU];
                T v[kSweepU], xv[kSweepU];
                if (k == 0) {
                    sweep_decode<CODED>(na.raw, cj, live, 0, len, i, n, dict);
#pragma unroll
                    for (int u = 0; u < kSweepU; ++u) v[u] = na.v[u];
                } else {
                    sweep_cols<CODED>(cj, live, k, L, len, base, wbase, lane, i, n, scol, code, dict);
#pragma unroll
                    for (int u = 0; u < kSweepU; ++u)   // slots past the slice width: not loaded (wave-uniform)
                        v[u] = k + u < L ? __builtin_nontemporal_load(&sval[base + (k + u) * 64 + lane])
                                         : static_cast<T>(0);
                }
#pragma unroll
                for (int u = 0; u < kSweepU; ++u) xv[u] = k + u < L ? out[cj[u]] : static_cast<T>(0);
                if (k + kSweepU >= L && w + stride < nsl) ahead(w + stride);   // after this step's gathers
#pragma unroll
                for (int u = 0; u < kSweepU; ++u) {
                    if (!live[u]) continue;
                    while (cj[u] < blk_lo) {   // next lower colour block
                        chunk += res;
                        res = 0;
                        --blk;
                        blk_lo = co.off[blk];
                    }
                    res += v[u] * xv[u];
                }
            }
            if (L == 0 && w + stride < nsl) ahead(w + stride);   // an empty slice ran no step
            if (valid) {
                chunk += res;
                const T zi = chunk * idv[i];
                out[i] = zi;
                if (RZ) rz += static_cast<double>(rin[i]) * static_cast<double>(zi);
            }
        }
    }
    if (RZ) {
        __shared__ double sh[kBlock / 64];
        rz = block_sum(rz, sh);
        if (threadIdx.x == 0) rzp[blockIdx.x] = rz;
    }
}

// r.z of a CG step from the sweeps' partials: one workgroup sums the m
// partials in a fixed order into *out.
__global__ __launch_bounds__(kFinBlock) void k_rz_sum(const double *__restrict__ part, int m, double *out,
                                                      const int *__restrict__ done) {
    if (done && *done) return;
    double v[1];
    reduce_partials<1>(part, 0, m, v);
    if (threadIdx.x == 0) *out = v[0];
}

// sweep values from the CSR values through the position map (-1: padding)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sweep_fill(long long m, const int *__restrict__ pos,
                                                       const T *__restrict__ val, T *__restrict__ out) {
    const long long k = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x;
    if (k >= m) return;
    const int j = pos[k];
    out[k] = j >= 0 ? val[j] : static_cast<T>(0);
}

// Level-scheduled triangular sweeps (CPU_CSR_lMatrix::Psgauss_seidel,
// lmatrix_csr_cpu.cc:542-616; ilu_solve, lmatrix_csr_cpu.cc:2108-2171).
// mode 0: SGS forward (D+L) y = r; mode 1: SGS backward (D+U) z = D y;
// mode 2: ILU unit-lower forward; mode 3: ILU (D+U) backward.  One wavefront
// per slice of a level, lane = row: the lane's row index, entry count,
// diagonal, columns and values are addressed by (slice, lane) alone, so they
// arrive together; only the gathers of earlier levels' results wait.  Each
// product and difference is rounded separately, in stored order, as the CPU
// loop does.
constexpr int kTsweepU = 4;

template <typename T, int MODE>
__device__ __forceinline__ void tsweep_slice(int s, int lane, const int *__restrict__ ptr, const int *__restrict__ srow,
                                             const int *__restrict__ slen, const int *__restrict__ scol,
                                             const T *__restrict__ sval, const T *__restrict__ dval,
                                             const T *__restrict__ in, T *out) {
    const int p = s * 64 + lane;
    const int i = srow[p];
    if (i < 0) return;   // padding lane of a level's last slice
    const int base = ptr[s], L = (ptr[s + 1] - base) >> 6;
    const int len = slen[p];
    const T d = dval[p];
    T acc;
    if (MODE == 0 || MODE == 2) acc = in[i];
    else if (MODE == 1) acc = mul_rn(out[i], d);
    else acc = out[i];
    for (int k = 0; k < L; k += kTsweepU) {
        int c[kTsweepU];
        T v[kTsweepU], xv[kTsweepU];
#pragma unroll
        for (int u = 0; u < kTsweepU; ++u) {
            const int idx = base + min(k + u, L - 1) * 64 + lane;
            c[u] = scol[idx];
            v[u] = sval[idx];
        }
#pragma unroll
        for (int u = 0; u < kTsweepU; ++u) xv[u] = out[c[u]];
#pragma unroll
        for (int u = 0; u < kTsweepU; ++u)
            if (k + u < len) acc = add_rn(acc, -mul_rn(xv[u], v[u]));
    }
    if (MODE != 2) acc = acc / d;
    out[i] = acc;
}

// One level: slices [s0, s1).
template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void k_tsweep(int s0, int s1, const int *__restrict__ ptr,
                                                   const int *__restrict__ srow, const int *__restrict__ slen,
                                                   const int *__restrict__ scol, const T *__restrict__ sval,
                                                   const T *__restrict__ dval, const T *__restrict__ in, T *out,
                                                   const int *__restrict__ done) {
    if (done && *done) return;
    const int s = s0 + blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    if (s >= s1) return;
    tsweep_slice<T, MODE>(s, threadIdx.x & 63, ptr, srow, slen, scol, sval, dval, in, out);
}

// A run of consecutive small levels [l0, l1) in one workgroup: the barrier
// between levels makes each level's results visible to the next, so the
// run costs one launch instead of l1 - l0.
template <typename T, int MODE>
__global__ __launch_bounds__(kMergeBlock) void k_tsweep_run(const int *__restrict__ lslice, int l0, int l1,
                                                            const int *__restrict__ ptr, const int *__restrict__ srow,
                                                            const int *__restrict__ slen, const int *__restrict__ scol,
                                                            const T *__restrict__ sval, const T *__restrict__ dval,
                                                            const T *__restrict__ in, T *out,
                                                            const int *__restrict__ done) {
    if (done && *done) return;
    constexpr int W = kMergeBlock / 64;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    for (int l = l0; l < l1; ++l) {
        for (int s = lslice[l] + wave; s < lslice[l + 1]; s += W)
            tsweep_slice<T, MODE>(s, lane, ptr, srow, slen, scol, sval, dval, in, out);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Sync-free triangular sweeps: one launch per sweep instead of one per level.
// Waves take slices in level order from ticket counters; a lane needing the
// result of an earlier row polls that result itself.  One counter hands out
// a ticket about every 11 ns (same-address atomics serialise), which bounds
// a sweep of 32,768 slices at about 0.4 ms: the tickets are dealt over
// kSfClasses counters instead, each on its own 128-byte line, and the waves
// of workgroup g take the tickets t = g mod kSfClasses (mod kSfClasses) in
// order from counter g mod kSfClasses.  Every result starts as
// a signalling-NaN sentinel (k_sf_fill, before the forward sweep).
// Arithmetic never yields a signalling NaN, so a value other than the
// sentinel is final.  Results are stored and polled with agent-scope relaxed
// atomics: coherent across the XCDs' L2 caches, with no fence.  A wave waits
// only on slices with smaller tickets, so the smallest unfinished slice
// always progresses: either a wave holds it, or it is the next ticket of its
// class, whose waves have finished every earlier (smaller) ticket of the
// class and take it next.  Every class has waves: the grid is a multiple of
// kSfClasses workgroups, few enough to be resident together.  Every wait is
// bounded: after kSfSpins polls, or once any lane has given up, a lane
// raises the failure word and goes on with the sentinel (the NaN then spreads
// through the results), so every wave ends and the drivers report the
// failure.  Each row's arithmetic is tsweep_slice's: results are
// bit-identical to the level launches.
template <typename T> struct Sent;
template <> struct Sent<double> {
    typedef unsigned long long U;
    static constexpr U bits = 0x7FF4DEAD5EEDBEEFull;   // exponent all ones, quiet bit clear
    static constexpr U quiet = 1ull << 51;
};
template <> struct Sent<float> {
    typedef unsigned int U;
    static constexpr U bits = 0x7FA5EED5u;
    static constexpr U quiet = 1u << 22;
};
constexpr int kSfSpins = 1 << 20;
constexpr int kSfClasses = 8;
// ctl: [2] failure word, [3] the most polls any lane needed for one result;
// ticket counter c of the forward (backward) sweep at int 32 (1 + c)
// (32 (1 + kSfClasses + c))
constexpr int kSfCtlInts = 32 * (1 + 2 * kSfClasses);

__device__ __forceinline__ int *sf_ticket(int *ctl, bool backward) {
    return ctl + 32 * (1 + (backward ? kSfClasses : 0) + static_cast<int>(blockIdx.x % kSfClasses));
}

// The next ticket of this workgroup's class (wave-uniform).
__device__ __forceinline__ int sf_next(int *counter, int lane) {
    int t = 0;
    if (lane == 0) t = __hip_atomic_fetch_add(counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __builtin_amdgcn_readfirstlane(t) * kSfClasses + static_cast<int>(blockIdx.x % kSfClasses);
}

template <typename T>
__device__ __forceinline__ typename Sent<T>::U sf_load(const T *p) {
    return __hip_atomic_load(reinterpret_cast<const typename Sent<T>::U *>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

// Stores a final result.  Only a copied input that holds the sentinel's own
// bits could equal it (arithmetic quiets every NaN); it is stored quieted.
template <typename T>
__device__ __forceinline__ void sf_store(T *p, T v) {
    typedef typename Sent<T>::U U;
    U u = __builtin_bit_cast(U, v);
    if (u == Sent<T>::bits) u |= Sent<T>::quiet;
    __hip_atomic_store(reinterpret_cast<U *>(p), u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits until none of the first m of w[] (results at out + c[]) is the
// sentinel, polling the pending ones together (bounded, see above); `most`
// keeps the largest poll count.
template <typename T, int U>
__device__ __forceinline__ void sf_wait_all(const T *out, const int *c, typename Sent<T>::U *w, int m, int *fail,
                                            int &most) {
    auto pending = [&]() {
        bool p = false;
#pragma unroll
        for (int u = 0; u < U; ++u) p |= u < m && w[u] == Sent<T>::bits;
        return p;
    };
    int spins = 0;
    while (pending()) {
        ++spins;
        if ((spins & 255) == 0 &&
            (spins >= kSfSpins || __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
            __hip_atomic_store(fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (u < m && w[u] == Sent<T>::bits) w[u] = sf_load(out + c[u]);
    }
    most = max(most, spins);
}

// y and z to the sentinel, every ticket counter to 0.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sf_fill(int n, T *y, T *z, int *ctl, const int *__restrict__ done) {
    if (done && *done) return;
    typedef typename Sent<T>::U U;
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * kBlock) {
        reinterpret_cast<U *>(y)[i] = Sent<T>::bits;
        reinterpret_cast<U *>(z)[i] = Sent<T>::bits;
    }
    if (blockIdx.x == 0 && threadIdx.x < 2 * kSfClasses) ctl[32 * (1 + threadIdx.x)] = 0;
}

// One slice of a sync-free sweep; MODE as tsweep_slice.  Forward (0, 2):
// out = y, read from `in`.  Backward (1, 3): out = z, y_i from `yin`.
template <typename T, int MODE>
__device__ __forceinline__ void sf_slice(int s, int lane, const int *__restrict__ ptr, const int *__restrict__ srow,
                                         const int *__restrict__ slen, const int *__restrict__ scol,
                                         const T *__restrict__ sval, const T *__restrict__ dval,
                                         const T *__restrict__ in, const T *__restrict__ yin, T *out, int *fail,
                                         int &most) {
    typedef typename Sent<T>::U U;
    const int p = s * 64 + lane;
    const int base = ptr[s], L = (ptr[s + 1] - base) >> 6;
    const int i = srow[p];
    if (i < 0) return;   // padding lane of a level's last slice
    const int len = slen[p];
    const T d = dval[p];
    T acc;
    if (MODE == 0 || MODE == 2) acc = in[i];
    else if (MODE == 1) acc = mul_rn(yin[i], d);
    else acc = yin[i];
    for (int k = 0; k < L; k += kTsweepU) {
        int c[kTsweepU];
        T v[kTsweepU];
        U w[kTsweepU];
#pragma unroll
        for (int u = 0; u < kTsweepU; ++u) {
            const int idx = base + min(k + u, L - 1) * 64 + lane;
            c[u] = scol[idx];
            v[u] = sval[idx];
        }
#pragma unroll
        for (int u = 0; u < kTsweepU; ++u) w[u] = k + u < len ? sf_load(out + c[u]) : U(0);
        sf_wait_all<T, kTsweepU>(out, c, w, len - k, fail, most);
#pragma unroll
        for (int u = 0; u < kTsweepU; ++u)
            if (k + u < len) acc = add_rn(acc, -mul_rn(__builtin_bit_cast(T, w[u]), v[u]));
    }
    if (MODE != 2) acc = acc / d;
    sf_store(out + i, acc);
}

// A whole sweep: each wave takes one slice per ticket until every slice is
// taken (ctl: see kSfCtlInts).
template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void k_tsweep_sf(int nslice, const int *__restrict__ ptr,
                                                      const int *__restrict__ srow, const int *__restrict__ slen,
                                                      const int *__restrict__ scol, const T *__restrict__ sval,
                                                      const T *__restrict__ dval, const T *__restrict__ in,
                                                      const T *__restrict__ yin, T *out, int *ctl,
                                                      const int *__restrict__ done) {
    if (done && *done) return;
    const int lane = threadIdx.x & 63;
    int *ticket = sf_ticket(ctl, MODE == 1 || MODE == 3);
    int most = 0;
    for (;;) {
        const int t = sf_next(ticket, lane);
        if (t >= nslice) break;
        sf_slice<T, MODE>(t, lane, ptr, srow, slen, scol, sval, dval, in, yin, out, ctl + 2, most);
    }
    most = wave_max(most);
    if (lane == 0 && most > 0) __hip_atomic_fetch_max(ctl + 3, most, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sync-free sweep over geometric blocks of a structured grid (4x4x4 or
// 8x8 cells, at most 64 rows, one wave each), taken in block-level order
// from the ticket counters.  A wave first waits for every result its block
// needs from other blocks (bounded, as sf_slice), then runs the block's own
// levels: at each, the rows of that level read the results of their own
// block from the wave's LDS slots (written at earlier levels by the lanes
// that computed them; a wave's LDS accesses complete in issue order), finish,
// and write their own.  One memory round trip per block instead of one per
// row level: a 128^3 7-point sweep has 94 block levels against 382 row
// levels.  Each row's arithmetic is sf_slice's (the starting value, the
// entries in stored order, then the diagonal): results are bit-identical.
// K bounds the entries of any lane.
template <typename T, int MODE, int K>
__global__ __launch_bounds__(kBlock) void k_bsweep_sf(int nblk, const int *__restrict__ ptr,
                                                      const int *__restrict__ brow, const int *__restrict__ blen,
                                                      const int *__restrict__ blev, const int *__restrict__ bnl,
                                                      const int *__restrict__ bcol, const T *__restrict__ bval,
                                                      const T *__restrict__ dval, const T *__restrict__ in,
                                                      const T *__restrict__ yin, T *out, int *ctl,
                                                      const int *__restrict__ done) {
    if (done && *done) return;
    typedef typename Sent<T>::U U;
    __shared__ T xch[kBlock];   // each lane's result, for the later rows of its block
    const int lane = threadIdx.x & 63;
    T *wx = xch + (threadIdx.x & ~63);
    int *ticket = sf_ticket(ctl, MODE == 1 || MODE == 3);
    int most = 0;
    for (;;) {
        const int b = sf_next(ticket, lane);
        if (b >= nblk) break;
        const int p = b * 64 + lane;
        const int base = ptr[b], L = (ptr[b + 1] - base) >> 6, nl = bnl[b];
        const int i = brow[p], len = blen[p], my = blev[p];   // padding lane: -1, 0, -1
        int c[K];
        T v[K];
        U w[K];
#pragma unroll
        for (int u = 0; u < K; ++u) {
            c[u] = 0;
            v[u] = T(0);
            if (u < L) {
                c[u] = bcol[base + u * 64 + lane];
                v[u] = bval[base + u * 64 + lane];
            }
        }
        // the row's own operands before the wait, off the critical path
        T d = T(1), acc0 = T(0);
        if (i >= 0) {
            d = dval[p];
            acc0 = MODE == 0 || MODE == 2 ? in[i] : yin[i];
        }
#pragma unroll
        for (int u = 0; u < K; ++u) w[u] = u < len && c[u] >= 0 ? sf_load(out + c[u]) : U(0);
        // only results from other blocks can still be the sentinel
        sf_wait_all<T, K>(out, c, w, len, ctl + 2, most);
        if (MODE == 1) acc0 = mul_rn(acc0, d);
        for (int l = 0; l < nl; ++l) {
            __builtin_amdgcn_wave_barrier();
            if (l == my) {
                // straight-line: every lane reads a slot for every entry (its
                // own for an entry from another block or a padding one), all
                // reads issued before the first use, then selects.  Padding
                // entries (u >= len) hold value 0 and operand +0, and
                // acc + -(0 * 0) = acc exactly, so they need no guard.
                T x[K];
#pragma unroll
                for (int u = 0; u < K; ++u) x[u] = wx[c[u] < 0 ? -1 - c[u] : lane];
#pragma unroll
                for (int u = 0; u < K; ++u)
                    if (c[u] >= 0) x[u] = __builtin_bit_cast(T, w[u]);
                T acc = acc0;
#pragma unroll
                for (int u = 0; u < K; ++u) acc = add_rn(acc, -mul_rn(x[u], v[u]));
                if (MODE != 2) acc = acc / d;
                wx[lane] = acc;
                sf_store(out + i, acc);
            }
        }
        __builtin_amdgcn_wave_barrier();   // the block's reads before the next block's writes
    }
    most = wave_max(most);
    if (lane == 0 && most > 0) __hip_atomic_fetch_max(ctl + 3, most, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ILU(0) of row i in place (CPU_CSR_lMatrix::ilu0, lmatrix_csr_cpu.cc:1880-1930,
// IKJ): for each lower entry a_ik (k < i, ascending) with a stored diagonal
// a_kk: a_ik /= a_kk, then a_ij -= a_ik a_kj for every later entry j of row i
// that row k also holds.  Rows k are final (earlier levels), and row i is
// written by its own thread only.  Products and differences are rounded
// separately, as in the CPU loop.
template <typename T>
__device__ __forceinline__ void ilu0_row(int i, const int *__restrict__ row, const int *__restrict__ col, T *val,
                                         const int *__restrict__ diag) {
#pragma clang fp contract(off)
    const int ie = row[i + 1];
    for (int ik = row[i]; ik < ie; ++ik) {
        const int k = col[ik];
        if (k >= i) break;
        const int dk = diag[k];
        if (dk < 0) continue;   // missing diagonal: counted on the host, entry left as is
        const T lik = val[ik] / val[dk];
        val[ik] = lik;
        int q = row[k];
        const int qe = row[k + 1];
        for (int ij = ik + 1; ij < ie; ++ij) {
            const int cj = col[ij];
            while (q < qe && col[q] < cj) ++q;
            if (q < qe && col[q] == cj) {
                const T prod = lik * val[q];
                val[ij] = val[ij] - prod;
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_ilu0_level(const int *__restrict__ lrows, int cnt,
                                                       const int *__restrict__ row, const int *__restrict__ col,
                                                       T *val, const int *__restrict__ diag) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= cnt) return;
    ilu0_row<T>(lrows[k], row, col, val, diag);
}

template <typename T>
__global__ __launch_bounds__(kMergeBlock) void k_ilu0_run(const int *__restrict__ lrows, const int *__restrict__ lptr,
                                                          int l0, int l1, const int *__restrict__ row,
                                                          const int *__restrict__ col, T *val,
                                                          const int *__restrict__ diag) {
    for (int l = l0; l < l1; ++l) {
        const int a = lptr[l], cnt = lptr[l + 1] - a;
        for (int k = threadIdx.x; k < cnt; k += kMergeBlock) ilu0_row<T>(lrows[a + k], row, col, val, diag);
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_mc_diag_setup(int n, const int *__restrict__ diag,
                                                          const T *__restrict__ val, T omega_relax,
                                                          T *__restrict__ idv, T *__restrict__ dv,
                                                          int *__restrict__ bad) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int p = diag[i];
    const T d = p >= 0 ? val[p] : static_cast<T>(0);
    if (d == static_cast<T>(0)) atomicAdd(bad, 1);
    // iDv: 1/d scaled by -1; Dv: d scaled by -1 then by 1/(omega(2-omega))
    idv[i] = (static_cast<T>(1) / d) * static_cast<T>(-1);
    dv[i] = (d * static_cast<T>(-1)) * omega_relax;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale_offdiag(int n, const int *__restrict__ row, const int *__restrict__ col,
                                                          const T *__restrict__ val, T omega, T *__restrict__ sval) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    for (int j = row[i]; j < row[i + 1]; ++j) sval[j] = col[j] == i ? val[j] : val[j] * omega;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_check_diag(int n, const int *__restrict__ diag, const T *__restrict__ val,
                                                       int *__restrict__ bad) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int p = diag[i];
    if (p < 0 || val[p] == static_cast<T>(0)) atomicAdd(bad, 1);
}

static int count_bad_diag(const hfh_csr *A, const char *what, void *idv, void *dv, double relax) {
    hfh_ctx *c = A->ctx;
    int *bad = nullptr;
    int rc = dev_alloc(reinterpret_cast<void **>(&bad), sizeof(int));
    if (rc) return rc;
    HFH_HIP(hipMemsetAsync(bad, 0, sizeof(int), c->stream));
    const int nb = (A->nrow + kBlock - 1) / kBlock;
    if (A->nrow > 0) {
        if (idv) {
            if (A->dtype == HFH_F64)
                hipLaunchKernelGGL(k_mc_diag_setup<double>, dim3(nb), dim3(kBlock), 0, c->stream, A->nrow, A->diag,
                                   (const double *)A->val, relax, (double *)idv, (double *)dv, bad);
            else
                hipLaunchKernelGGL(k_mc_diag_setup<float>, dim3(nb), dim3(kBlock), 0, c->stream, A->nrow, A->diag,
                                   (const float *)A->val, (float)relax, (float *)idv, (float *)dv, bad);
        } else {
            if (A->dtype == HFH_F64)
                hipLaunchKernelGGL(k_check_diag<double>, dim3(nb), dim3(kBlock), 0, c->stream, A->nrow, A->diag, (const double *)A->val, bad);
            else
                hipLaunchKernelGGL(k_check_diag<float>, dim3(nb), dim3(kBlock), 0, c->stream, A->nrow, A->diag, (const float *)A->val, bad);
        }
        if ((rc = launch_check(what))) { (void)hipFree(bad); return rc; }
    }
    int hbad = 0;
    HFH_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HFH_HIP(hipStreamSynchronize(c->stream));
    (void)hipFree(bad);
    HFH_CHECK(hbad == 0, HFH_ERR_ZERO_PIVOT, "%s: %d rows with a zero or missing diagonal entry", what, hbad);
    return HFH_OK;
}

static int upload_levels(hfh_levels &L, int n, int nlev, const int *ptr, const int *rows, const char *what) {
    HFH_CHECK(nlev >= 0 && ptr && rows, HFH_ERR_INVALID, "%s: missing level schedule", what);
    HFH_CHECK(ptr[0] == 0 && ptr[nlev] == n, HFH_ERR_INVALID, "%s: level pointer must run from 0 to %d", what, n);
    std::vector<char> seen(n, 0);
    for (int l = 0; l < nlev; ++l) HFH_CHECK(ptr[l + 1] >= ptr[l], HFH_ERR_INVALID, "%s: level pointer decreases", what);
    for (int k = 0; k < n; ++k) {
        HFH_CHECK(rows[k] >= 0 && rows[k] < n && !seen[rows[k]], HFH_ERR_INVALID,
                  "%s: level rows are not a permutation (entry %d)", what, k);
        seen[rows[k]] = 1;
    }
    L.nlev = nlev;
    L.ptr.assign(ptr, ptr + nlev + 1);
    int rc = dev_alloc(reinterpret_cast<void **>(&L.rows), sizeof(int) * static_cast<size_t>(n));
    if (!rc) rc = dev_alloc(reinterpret_cast<void **>(&L.ptr_d), sizeof(int) * static_cast<size_t>(nlev + 1));
    if (rc) return rc;
    HFH_HIP(hipMemcpy(L.rows, rows, sizeof(int) * static_cast<size_t>(n), hipMemcpyHostToDevice));
    HFH_HIP(hipMemcpy(L.ptr_d, ptr, sizeof(int) * static_cast<size_t>(nlev + 1), hipMemcpyHostToDevice));
    // launch plan: runs of at least two consecutive levels of at most
    // kMergeBlock rows share one launch; other levels get one launch each
    L.runs.clear();
    for (int l = 0; l < nlev;) {
        int e = l;
        while (e < nlev && ptr[e + 1] - ptr[e] <= kMergeBlock) ++e;
        if (e - l >= 2) {
            L.runs.emplace_back(l, e);
            l = e;
        } else {
            L.runs.emplace_back(l, l + 1);
            ++l;
        }
    }
    return HFH_OK;
}

// Checks that every lower (forward) / upper (backward) dependency of a row sits
// in an earlier level, using the host copy of the structure.
static int validate_levels(const hfh_csr *A, const hfh_levels &L, bool upper, const char *what) {
    const int n = A->nrow;
    std::vector<int> col(A->nnz), lev(n), rows(n);
    HFH_HIP(hipMemcpy(col.data(), A->col, sizeof(int) * static_cast<size_t>(A->nnz), hipMemcpyDeviceToHost));
    HFH_HIP(hipMemcpy(rows.data(), L.rows, sizeof(int) * static_cast<size_t>(n), hipMemcpyDeviceToHost));
    for (int l = 0; l < L.nlev; ++l)
        for (int k = L.ptr[l]; k < L.ptr[l + 1]; ++k) lev[rows[k]] = l;
    for (int i = 0; i < n; ++i)
        for (int j = A->h_row[i]; j < A->h_row[i + 1]; ++j) {
            const int c = col[j];
            const bool dep = upper ? c > i : c < i;
            HFH_CHECK(!dep || lev[c] < lev[i], HFH_ERR_INVALID,
                      "%s: row %d (level %d) depends on row %d (level %d) in the %s sweep", what, i, lev[i], c,
                      lev[c], upper ? "backward" : "forward");
        }
    return HFH_OK;
}

template <typename T, int MODE>
static void sweep(const hfh_levels &Lv, const hfh_tsweep &W, const T *in, T *out, const int *done, hipStream_t s) {
    const T *sv = static_cast<const T *>(W.val), *dv = static_cast<const T *>(W.dval);
    for (const auto &run : Lv.runs) {
        if (run.second - run.first > 1) {
            hipLaunchKernelGGL((k_tsweep_run<T, MODE>), dim3(1), dim3(kMergeBlock), 0, s, W.lslice_d, run.first,
                               run.second, W.ptr, W.row, W.len, W.col, sv, dv, in, out, done);
            continue;
        }
        const int s0 = W.lslice[run.first], s1 = W.lslice[run.first + 1];
        if (s1 == s0) continue;
        constexpr int WPB = kBlock / 64;
        hipLaunchKernelGGL((k_tsweep<T, MODE>), dim3((s1 - s0 + WPB - 1) / WPB), dim3(kBlock), 0, s, s0, s1, W.ptr,
                           W.row, W.len, W.col, sv, dv, in, out, done);
    }
}

// One blocked sweep, its entry bound K from the most entries of any lane.
template <typename T, int MODE>
static void bsweep_launch(const hfh_tsweep &W, int grid, const T *in, const T *yin, T *out, int *ctl,
                          const int *done, hipStream_t s) {
#define HFH_BSWEEP(K)                                                                                             \
    hipLaunchKernelGGL((k_bsweep_sf<T, MODE, K>), dim3(grid), dim3(kBlock), 0, s, W.nslice, W.ptr, W.row, W.len, \
                       W.lev, W.bnl, W.col, (const T *)W.val, (const T *)W.dval, in, yin, out, ctl, done)
    if (W.kmax <= 3) HFH_BSWEEP(3);   // 7-point stencils
    else if (W.kmax <= 4) HFH_BSWEEP(4);
    else if (W.kmax <= 8) HFH_BSWEEP(8);
    else HFH_BSWEEP(16);
#undef HFH_BSWEEP
}

// Sync-free form of both sweeps: the fill, the forward sweep into P->sf_y,
// the backward sweep into out.
template <typename T, int FW, int BW>
static void sf_sweeps(hfh_prec *P, const T *in, T *out, const int *done) {
    hipStream_t s = P->ctx->stream;
    T *y = static_cast<T *>(P->sf_y);
    const int nfill = std::max(1, std::min((P->n + kBlock - 1) / kBlock, std::max(P->ctx->num_cu, 1) * 8));
    hipLaunchKernelGGL(k_sf_fill<T>, dim3(nfill), dim3(kBlock), 0, s, P->n, y, out, P->sf_ctl, done);
    // Waiting waves poll, and every poll is a memory round trip: more waves
    // than the levels keep busy slow all of them (upwind 128^3 ILU(0) apply:
    // 1.40 ms at one wave per CU, 1.85 ms at 4, 2.39 ms at 6, and at 32 the
    // polls flood the memory system).  Automatic (sf_waves = 0): about three
    // levels' slices in flight, between 1 and 8 waves per CU.
    const int ncu = std::max(P->ctx->num_cu, 1);
    auto grid = [&](const hfh_tsweep &W) {
        int waves;
        if (P->sf_waves > 0) {
            waves = ncu * P->sf_waves;
        } else {
            const int nlev = std::max<int>(1, static_cast<int>(W.lslice.size()) - 1);
            waves = std::min(ncu * 8, std::max(ncu, 3 * ((W.nslice + nlev - 1) / nlev)));
        }
        waves = std::min(waves, W.nslice);
        const int wg = (waves + kBlock / 64 - 1) / (kBlock / 64);
        return std::max(1, (wg + kSfClasses - 1) / kSfClasses) * kSfClasses;   // every class has waves
    };
    // geometric blocks where the grid was recognised (HFH_SWEEP_SYNCFREE),
    // else slices of one level
    const bool blocked = P->sweep_mode == HFH_SWEEP_SYNCFREE;
    const hfh_tsweep &L = blocked && P->blo.nslice > 0 ? P->blo : P->tlo;
    const hfh_tsweep &U = blocked && P->bup.nslice > 0 ? P->bup : P->tup;
    if (L.nslice > 0) {
        if (L.lev) bsweep_launch<T, FW>(L, grid(L), in, nullptr, y, P->sf_ctl, done, s);
        else
            hipLaunchKernelGGL((k_tsweep_sf<T, FW>), dim3(grid(L)), dim3(kBlock), 0, s, L.nslice, L.ptr, L.row,
                               L.len, L.col, (const T *)L.val, (const T *)L.dval, in, (const T *)nullptr, y,
                               P->sf_ctl, done);
    }
    if (U.nslice > 0) {
        if (U.lev) bsweep_launch<T, BW>(U, grid(U), nullptr, y, out, P->sf_ctl, done, s);
        else
            hipLaunchKernelGGL((k_tsweep_sf<T, BW>), dim3(grid(U)), dim3(kBlock), 0, s, U.nslice, U.ptr, U.row,
                               U.len, U.col, (const T *)U.val, (const T *)U.dval, (const T *)nullptr, (const T *)y,
                               out, P->sf_ctl, done);
    }
}

template <typename T>
static int trsv_sweeps(hfh_prec *P, const T *in, T *out, const int *done) {
    hipStream_t s = P->ctx->stream;
    if (P->sweep_mode != HFH_SWEEP_LEVELS && P->sf_y && in != out) {
        if (P->kind == HFH_PREC_ILU0) sf_sweeps<T, 2, 3>(P, in, out, done);
        else sf_sweeps<T, 0, 1>(P, in, out, done);
        return launch_check("k_tsweep_sf");
    }
    if (P->kind == HFH_PREC_ILU0) {
        sweep<T, 2>(P->lo, P->tlo, in, out, done, s);
        sweep<T, 3>(P->up, P->tup, in, out, done, s);
    } else {
        sweep<T, 0>(P->lo, P->tlo, in, out, done, s);
        sweep<T, 1>(P->up, P->tup, in, out, done, s);
    }
    return launch_check("k_tsweep");
}

// r.z partials that a multicolour SGS apply with rzp stores: one per
// workgroup of the steps that finalise z (an upper bound for any colouring).
int mcsgs_rz_slots(const hfh_prec *P) {
    constexpr int WPB = kBlock / 64;
    int m = 0;
    for (int c = 0; c < P->ncolors; ++c) {
        const int nf = P->fw.slice0[c + 1] - P->fw.slice0[c], nb = P->bw.slice0[c + 1] - P->bw.slice0[c];
        m += (std::max(nf, nb) + WPB - 1) / WPB;
    }
    return std::max(m, 1);
}

template <typename T>
static int mcsgs_apply(hfh_prec *P, const T *in, T *out, const int *done, bool first_done, double *rzp,
                       double *rz_out) {
    hipStream_t s = P->ctx->stream;
    ColorOffsets co;
    for (int c = 0; c <= P->ncolors; ++c) co.off[c] = P->color_off[c];
    const T *idv = (const T *)P->invd, *dv = (const T *)P->dv;
    const int nc = P->ncolors;
    constexpr int WPB = kBlock / 64;
    // with rzp, the steps that finalise z (the last colour's forward step,
    // every backward step) store r.z partials there, one per workgroup
    int nrz = 0;
    // workgroups of a colour step: one slice per wave, or at most
    // kSweepGridPerCu per CU with waves taking several slices
    const int gcap = std::max(P->ctx->num_cu, 1) * kSweepGridPerCu<T>;
    auto grid = [&](int nsl) {
        const int need = (nsl + WPB - 1) / WPB;
        return std::max(1, gcap > 0 ? std::min(need, gcap) : need);
    };
    auto fwd = [&](auto coded, auto rz, int c, int s0, int nsl) {
        const hfh_sweep &W = P->fw;
        const int g = grid(nsl);
        hipLaunchKernelGGL((k_mcs_fwd<T, decltype(coded)::value, decltype(rz)::value>), dim3(g), dim3(kBlock), 0, s, c,
                           co, c == nc - 1, s0, nsl, P->n, W.ptr, W.len, W.col, W.code, W.wptr, W.dict, W.ndict,
                           (const T *)W.val, in, idv, dv, out, done, rzp + nrz);
        if (decltype(rz)::value) nrz += g;
    };
    auto bwd = [&](auto coded, auto rz, int c, int s0, int nsl) {
        const hfh_sweep &W = P->bw;
        const int g = grid(nsl);
        hipLaunchKernelGGL((k_mcs_bwd<T, decltype(coded)::value, decltype(rz)::value>), dim3(g), dim3(kBlock), 0, s, c,
                           nc, co, s0, nsl, P->n, W.ptr, W.len, W.col, W.code, W.wptr, W.dict, W.ndict,
                           (const T *)W.val, idv, dv, out, done, in, rzp + nrz);
        if (decltype(rz)::value) nrz += g;
    };
    if (rzp) {
        int need = 0;
        for (int c = 0; c < nc; ++c) {
            if (co.off[c + 1] == co.off[c]) continue;
            const hfh_sweep &W = c == nc - 1 ? P->fw : P->bw;
            need += grid(W.slice0[c + 1] - W.slice0[c]);
        }
        if (need > mcsgs_rz_slots(P)) {
            set_error("mcsgs_apply: %d r.z partials exceed the %d slots", need, mcsgs_rz_slots(P));
            return HFH_ERR_INVALID;
        }
    }
    using Yes = std::true_type;
    using No = std::false_type;
    for (int c = 0; c < nc; ++c) {
        const int cnt = co.off[c + 1] - co.off[c];
        if (cnt == 0) continue;
        if (c == 0 && nc > 1) {
            if (first_done) continue;
            hipLaunchKernelGGL(k_mc_fwd0<T>, dim3((cnt + kBlock - 1) / kBlock), dim3(kBlock), 0, s, co.off[0],
                               co.off[1], in, idv, out, done);
        } else {
            const int s0 = P->fw.slice0[c], nsl = P->fw.slice0[c + 1] - s0;
            const bool rz = rzp && c == nc - 1;
            if (P->fw.code)
                rz ? fwd(Yes{}, Yes{}, c, s0, nsl) : fwd(Yes{}, No{}, c, s0, nsl);
            else
                rz ? fwd(No{}, Yes{}, c, s0, nsl) : fwd(No{}, No{}, c, s0, nsl);
        }
    }
    for (int c = nc - 2; c >= 0; --c) {
        const int cnt = co.off[c + 1] - co.off[c];
        if (cnt == 0) continue;
        const int s0 = P->bw.slice0[c], nsl = P->bw.slice0[c + 1] - s0;
        if (P->bw.code)
            rzp ? bwd(Yes{}, Yes{}, c, s0, nsl) : bwd(Yes{}, No{}, c, s0, nsl);
        else
            rzp ? bwd(No{}, Yes{}, c, s0, nsl) : bwd(No{}, No{}, c, s0, nsl);
    }
    if (rzp)
        hipLaunchKernelGGL(k_rz_sum, dim3(1), dim3(kFinBlock), 0, s, (const double *)rzp, nrz, rz_out, done);
    return launch_check("k_mc_sweep");
}

// Builds one sweep copy from the host CSR structure.  forward: entries with a
// column in a lower colour (CSR order); backward: entries in a higher colour,
// block by block from the highest colour down.
static int build_sweep(hfh_prec *P, const hfh_csr *A, bool forward, hfh_sweep &W) {
    const int nc = P->ncolors, n = A->nrow;
    const std::vector<int> &off = P->color_off;
    const std::vector<int> &row = A->h_row, &col = A->h_col;
    std::vector<int> len(std::max(n, 1), 0);
    W.slice0.assign(nc + 1, 0);
    for (int c = 0; c < nc; ++c) W.slice0[c + 1] = W.slice0[c] + (off[c + 1] - off[c] + 63) / 64;
    const int nsl = W.slice0[nc];
    // entries of row i in sweep order
    std::vector<std::pair<int, int>> segs;
    auto entries = [&](int c, int i, std::vector<int> &e) {
        e.clear();
        if (forward) {
            for (int j = row[i]; j < row[i + 1] && col[j] < off[c]; ++j) e.push_back(j);
            return;
        }
        // columns ascend, so the higher-colour entries split into one
        // segment per colour block; emit the segments from the highest down
        segs.clear();
        int j = row[i];
        const int je = row[i + 1];
        while (j < je && col[j] < off[c + 1]) ++j;
        int b = c + 1;
        while (j < je) {
            while (col[j] >= off[b + 1]) ++b;
            const int st = j;
            while (j < je && col[j] < off[b + 1]) ++j;
            segs.emplace_back(st, j);
        }
        for (auto it = segs.rbegin(); it != segs.rend(); ++it)
            for (int k = it->first; k < it->second; ++k) e.push_back(k);
    };
    std::vector<int> ptr(nsl + 1, 0), e, width(nsl, 0);
    // the distinct offsets col - row of the entries (0 included: padding)
    std::vector<long long> offs(1, 0);
    bool coded = true;
    for (int c = 0; c < nc; ++c)
        for (int w = 0; w < W.slice0[c + 1] - W.slice0[c]; ++w) {
            int L = 0;
            for (int lane = 0; lane < 64; ++lane) {
                const int i = off[c] + w * 64 + lane;
                if (i >= off[c + 1]) break;
                entries(c, i, e);
                len[i] = static_cast<int>(e.size());
                L = std::max(L, len[i]);
                for (int k = 0; coded && k < len[i]; ++k) {
                    const long long o = static_cast<long long>(col[e[k]]) - i;
                    auto it = std::lower_bound(offs.begin(), offs.end(), o);
                    if (it == offs.end() || *it != o) {
                        offs.insert(it, o);
                        coded = offs.size() <= kPadCode;   // code kPadCode marks padding slots
                    }
                }
            }
            width[W.slice0[c] + w] = L;
        }
    // values keep each slice's width; coded: the code words cover it rounded
    // up to whole words of 4 slots (wptr: word offset of each slice)
    long long total = 0, wtotal = 0;
    std::vector<int> wptr(nsl + 1, 0);
    for (int sl = 0; sl < nsl; ++sl) {
        ptr[sl] = static_cast<int>(total);
        wptr[sl] = static_cast<int>(wtotal);
        total += 64LL * width[sl];
        wtotal += 64LL * ((width[sl] + 3) / 4);
        HFH_CHECK(total < (1LL << 31), HFH_ERR_UNSUPPORTED, "hfh_mcsgs_create: sweep copy exceeds 2^31 entries");
    }
    ptr[nsl] = static_cast<int>(total);
    wptr[nsl] = static_cast<int>(wtotal);
    std::vector<int> pcol(coded ? 1 : std::max<long long>(total, 1)), pos(std::max<long long>(total, 1));
    std::vector<unsigned> words(coded ? std::max<long long>(wtotal, 1) : 1, 0u);
    for (int c = 0; c < nc; ++c)
        for (int w = 0; w < W.slice0[c + 1] - W.slice0[c]; ++w) {
            const int sl = W.slice0[c] + w;
            const int base = ptr[sl], L = (ptr[sl + 1] - base) / 64, wbase = wptr[sl];
            for (int lane = 0; lane < 64; ++lane) {
                const int i = off[c] + w * 64 + lane;
                const bool valid = i < off[c + 1];
                if (valid) entries(c, i, e);
                else e.clear();
                const int own = std::min(i, off[c + 1] - 1);   // a row of this slice
                for (int k = 0; k < (coded ? (L + 3) / 4 * 4 : L); ++k) {
                    const int idx = base + k * 64 + lane;
                    const bool real = k < static_cast<int>(e.size());
                    if (k < L) pos[idx] = real ? e[k] : -1;
                    if (!coded) {
                        pcol[idx] = real ? col[e[k]] : own;
                        continue;
                    }
                    const unsigned code =
                        real ? static_cast<unsigned>(std::lower_bound(offs.begin(), offs.end(),
                                                                      static_cast<long long>(col[e[k]]) - i) -
                                                     offs.begin())
                             : kPadCode;
                    words[wbase + (k / 4) * 64 + lane] |= code << (8 * (k % 4));
                }
            }
        }
    W.nslice = nsl;
    int rc;
    if ((rc = dev_alloc(reinterpret_cast<void **>(&W.ptr), sizeof(int) * ptr.size())) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&W.len), sizeof(int) * len.size())) ||
        (rc = dev_alloc(&W.val, A->elem() * pos.size())))
        return rc;
    HFH_HIP(hipMemcpy(W.ptr, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice));
    HFH_HIP(hipMemcpy(W.len, len.data(), sizeof(int) * len.size(), hipMemcpyHostToDevice));
    if (coded) {
        std::vector<int> dict(offs.begin(), offs.end());
        if ((rc = dev_alloc(reinterpret_cast<void **>(&W.code), sizeof(unsigned) * words.size())) ||
            (rc = dev_alloc(reinterpret_cast<void **>(&W.wptr), sizeof(int) * wptr.size())) ||
            (rc = dev_alloc(reinterpret_cast<void **>(&W.dict), sizeof(int) * dict.size())))
            return rc;
        HFH_HIP(hipMemcpy(W.code, words.data(), sizeof(unsigned) * words.size(), hipMemcpyHostToDevice));
        HFH_HIP(hipMemcpy(W.wptr, wptr.data(), sizeof(int) * wptr.size(), hipMemcpyHostToDevice));
        HFH_HIP(hipMemcpy(W.dict, dict.data(), sizeof(int) * dict.size(), hipMemcpyHostToDevice));
        W.ndict = static_cast<int>(dict.size());
    } else {
        if ((rc = dev_alloc(reinterpret_cast<void **>(&W.col), sizeof(int) * pcol.size()))) return rc;
        HFH_HIP(hipMemcpy(W.col, pcol.data(), sizeof(int) * pcol.size(), hipMemcpyHostToDevice));
    }
    int *dpos = nullptr;
    if ((rc = dev_alloc(reinterpret_cast<void **>(&dpos), sizeof(int) * pos.size()))) return rc;
    HFH_HIP(hipMemcpy(dpos, pos.data(), sizeof(int) * pos.size(), hipMemcpyHostToDevice));
    const long long m = static_cast<long long>(pos.size());
    const dim3 g(static_cast<unsigned>((m + kBlock - 1) / kBlock));
    const void *src = P->sval ? P->sval : A->val;
    hipStream_t s = A->ctx->stream;
    if (A->dtype == HFH_F64)
        hipLaunchKernelGGL(k_sweep_fill<double>, g, dim3(kBlock), 0, s, m, dpos, (const double *)src, (double *)W.val);
    else
        hipLaunchKernelGGL(k_sweep_fill<float>, g, dim3(kBlock), 0, s, m, dpos, (const float *)src, (float *)W.val);
    rc = launch_check("k_sweep_fill");
    (void)hipStreamSynchronize(s);
    (void)hipFree(dpos);
    return rc;
}

// Copies the values of a sliced triangular sweep from the CSR values
// through its position maps (padding: 0).
static int tsweep_fill(const hfh_csr *A, hfh_tsweep &W) {
    hipStream_t s = A->ctx->stream;
    const long long m = std::max(W.nent, 1LL), md = 64LL * std::max(W.nslice, 1);
    const dim3 g(static_cast<unsigned>((m + kBlock - 1) / kBlock)), gd(static_cast<unsigned>((md + kBlock - 1) / kBlock));
    if (A->dtype == HFH_F64) {
        hipLaunchKernelGGL(k_sweep_fill<double>, g, dim3(kBlock), 0, s, m, W.pos, (const double *)A->val, (double *)W.val);
        hipLaunchKernelGGL(k_sweep_fill<double>, gd, dim3(kBlock), 0, s, md, W.dpos, (const double *)A->val, (double *)W.dval);
    } else {
        hipLaunchKernelGGL(k_sweep_fill<float>, g, dim3(kBlock), 0, s, m, W.pos, (const float *)A->val, (float *)W.val);
        hipLaunchKernelGGL(k_sweep_fill<float>, gd, dim3(kBlock), 0, s, md, W.dpos, (const float *)A->val, (float *)W.dval);
    }
    return launch_check("k_sweep_fill");
}

// Builds the sliced copy of one sweep from the host structure and the level
// schedule (ptr, rows: host, level order): forward takes each row's lower
// entries, backward its upper ones, in stored order.
static int build_tsweep(const hfh_csr *A, const hfh_levels &Lv, const int *lrows, bool forward, hfh_tsweep &W) {
    const std::vector<int> &row = A->h_row, &col = A->h_col;
    const int nlev = Lv.nlev;
    W.lslice.assign(nlev + 1, 0);
    for (int l = 0; l < nlev; ++l) W.lslice[l + 1] = W.lslice[l] + (Lv.ptr[l + 1] - Lv.ptr[l] + 63) / 64;
    const int nsl = W.lslice[nlev];
    std::vector<int> sptr(nsl + 1, 0), srow(64LL * std::max(nsl, 1), -1), slen(64LL * std::max(nsl, 1), 0),
        sdpos(64LL * std::max(nsl, 1), -1);
    auto dep = [&](int c, int i) { return forward ? c < i : c > i; };
    long long total = 0;
    for (int l = 0; l < nlev; ++l)
        for (int sl = W.lslice[l]; sl < W.lslice[l + 1]; ++sl) {
            int L = 0;
            for (int lane = 0; lane < 64; ++lane) {
                const int k = Lv.ptr[l] + (sl - W.lslice[l]) * 64 + lane;
                if (k >= Lv.ptr[l + 1]) break;
                const int i = lrows[k], p = sl * 64 + lane;
                int cnt = 0;
                for (int j = row[i]; j < row[i + 1]; ++j) {
                    if (col[j] == i) sdpos[p] = j;
                    cnt += dep(col[j], i);
                }
                srow[p] = i;
                slen[p] = cnt;
                L = std::max(L, cnt);
            }
            sptr[sl] = static_cast<int>(total);
            total += 64LL * L;
            HFH_CHECK(total < (1LL << 31), HFH_ERR_UNSUPPORTED, "sweep copy exceeds 2^31 entries");
        }
    sptr[nsl] = static_cast<int>(total);
    std::vector<int> scol(std::max(total, 1LL), 0), spos(std::max(total, 1LL), -1);
    for (int sl = 0; sl < nsl; ++sl)
        for (int lane = 0; lane < 64; ++lane) {
            const int p = sl * 64 + lane, i = srow[p];
            if (i < 0) continue;
            int k = 0;
            for (int j = row[i]; j < row[i + 1]; ++j)
                if (dep(col[j], i)) {
                    const long long idx = sptr[sl] + 64LL * k + lane;
                    scol[idx] = col[j];
                    spos[idx] = j;
                    ++k;
                }
        }
    W.nslice = nsl;
    W.nent = total;
    int rc;
    if ((rc = dev_alloc(reinterpret_cast<void **>(&W.lslice_d), sizeof(int) * W.lslice.size())) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&W.ptr), sizeof(int) * sptr.size())) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&W.row), sizeof(int) * srow.size())) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&W.len), sizeof(int) * slen.size())) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&W.dpos), sizeof(int) * sdpos.size())) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&W.col), sizeof(int) * scol.size())) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&W.pos), sizeof(int) * spos.size())) ||
        (rc = dev_alloc(&W.val, A->elem() * scol.size())) ||
        (rc = dev_alloc(&W.dval, A->elem() * sdpos.size())))
        return rc;
    HFH_HIP(hipMemcpy(W.lslice_d, W.lslice.data(), sizeof(int) * W.lslice.size(), hipMemcpyHostToDevice));
    HFH_HIP(hipMemcpy(W.ptr, sptr.data(), sizeof(int) * sptr.size(), hipMemcpyHostToDevice));
    HFH_HIP(hipMemcpy(W.row, srow.data(), sizeof(int) * srow.size(), hipMemcpyHostToDevice));
    HFH_HIP(hipMemcpy(W.len, slen.data(), sizeof(int) * slen.size(), hipMemcpyHostToDevice));
    HFH_HIP(hipMemcpy(W.dpos, sdpos.data(), sizeof(int) * sdpos.size(), hipMemcpyHostToDevice));
    HFH_HIP(hipMemcpy(W.col, scol.data(), sizeof(int) * scol.size(), hipMemcpyHostToDevice));
    HFH_HIP(hipMemcpy(W.pos, spos.data(), sizeof(int) * spos.size(), hipMemcpyHostToDevice));
    return tsweep_fill(A, W);
}

// Geometry of a natural-order structured grid, read from the offsets i - c
// of the lower entries of a sample of rows (7-, 9-, 27-point stencils and
// the like): nx is the middle one of the smallest cluster of frequent
// offsets above 1, nxy that of the next cluster (0: a 2D grid).  False when
// the pattern shows no such grid.
static bool detect_grid(const hfh_csr *A, int &nx, long long &nxy) {
    const std::vector<int> &row = A->h_row, &col = A->h_col;
    const int n = A->nrow;
    if (n < 4096) return false;
    const int step = std::max(1, n / 65536);
    std::vector<long long> off;
    long long rows = 0;
    for (int i = 0; i < n; i += step, ++rows)
        for (int j = row[i]; j < row[i + 1]; ++j)
            if (col[j] < i - 1) off.push_back(static_cast<long long>(i) - col[j]);
    std::sort(off.begin(), off.end());
    std::vector<long long> freq;   // offsets in at least 1/8 of the sampled rows
    for (size_t a = 0; a < off.size();) {
        size_t e = a;
        while (e < off.size() && off[e] == off[a]) ++e;
        if (8LL * static_cast<long long>(e - a) >= rows) freq.push_back(off[a]);
        a = e;
    }
    if (freq.empty()) return false;
    size_t e1 = 0;
    while (e1 < freq.size() && freq[e1] < 2 * freq[0]) ++e1;
    nx = static_cast<int>(freq[(e1 - 1) / 2]);
    nxy = 0;
    if (nx < 8) return false;
    if (e1 < freq.size()) {
        size_t e2 = e1;
        while (e2 < freq.size() && freq[e2] <= freq[e1] + 2LL * nx + 2) ++e2;
        nxy = freq[e1 + (e2 - e1 - 1) / 2];
        if (nxy % nx != 0 || nxy / nx < 4) return false;
    }
    return true;
}

// Builds the blocked copy of one sweep (k_bsweep_sf) for an nx (x nxy / nx)
// grid: blocks of 4x4x4 cells (3D) or 8x8 (2D), each row's level inside its
// block from its dependencies there, the blocks ordered by their level in
// the graph of dependencies between blocks.  With `shear` the blocks tile
// the coordinates (x + y + 2z, y + z, z) instead: a 9- or 27-point stencil
// has a lower neighbour at (x + 1, y - 1), which makes side-by-side blocks
// wait for each other, while in those coordinates no lower neighbour lies
// ahead in any direction.  Leaves W empty (the slice form then runs) if the
// graph between blocks has a cycle or a row more than 16 entries.
static int build_bsweep(const hfh_csr *A, int nx, long long nxy, bool forward, bool shear, hfh_tsweep &W) {
    const std::vector<int> &row = A->h_row, &col = A->h_col;
    const int n = A->nrow;
    const int bx = nxy ? 4 : 8, by = bx, bz = nxy ? 4 : 1;
    const long long ny = nxy ? nxy / nx : (n + nx - 1) / nx;
    const long long nz = nxy ? (n + nxy - 1) / nxy : 1;
    const long long xr = shear ? nx + ny + 2 * nz : nx, yr = shear ? ny + nz : ny;
    const long long nbx = (xr + bx - 1) / bx, nby = (yr + by - 1) / by, nbz = (nz + bz - 1) / bz;
    if (nbx * nby * nbz >= (1LL << 30)) return HFH_OK;
    const int nraw = static_cast<int>(nbx * nby * nbz);
    std::vector<int> bid(n), bcnt(nraw + 1, 0);
    for (int i = 0; i < n; ++i) {
        long long x = i % nx, t = i / nx, y = t % ny, z = t / ny;
        if (shear) {
            x += y + 2 * z;
            y += z;
        }
        bid[i] = static_cast<int>((z / bz * nby + y / by) * nbx + x / bx);
        ++bcnt[bid[i] + 1];
    }
    for (int b = 0; b < nraw; ++b) bcnt[b + 1] += bcnt[b];
    // rows of each raw block, ascending; a row's lane is its place there
    std::vector<int> brows(n), lane(n), fill(bcnt.begin(), bcnt.end() - 1);
    for (int i = 0; i < n; ++i) {
        lane[i] = fill[bid[i]] - bcnt[bid[i]];
        brows[fill[bid[i]]++] = i;
    }
    auto dep = [&](int c, int i) { return forward ? c < i : c > i; };
    // level of each row inside its block (dependencies there come first in
    // row order, ascending forward, descending backward) and the edges
    // between blocks
    std::vector<int> ilev(n, 0);
    std::vector<std::pair<int, int>> edges;   // (block, block it waits for)
    int kmax = 0;
    for (int q = 0; q < n; ++q) {
        const int i = forward ? q : n - 1 - q;
        int l = 0, cnt = 0;
        for (int j = row[i]; j < row[i + 1]; ++j) {
            const int c = col[j];
            if (!dep(c, i)) continue;
            ++cnt;
            if (bid[c] == bid[i]) l = std::max(l, ilev[c] + 1);
            else edges.emplace_back(bid[i], bid[c]);
        }
        ilev[i] = l;
        kmax = std::max(kmax, cnt);
    }
    if (kmax > 16) return HFH_OK;
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    // block levels (Kahn): a block runs after every block it waits for
    std::vector<int> indeg(nraw, 0), sptr(nraw + 1, 0), succ(edges.size()), blev(nraw, 0);
    for (const auto &e : edges) {
        ++indeg[e.first];
        ++sptr[e.second + 1];
    }
    for (int b = 0; b < nraw; ++b) sptr[b + 1] += sptr[b];
    {
        std::vector<int> f(sptr.begin(), sptr.end() - 1);
        for (const auto &e : edges) succ[f[e.second]++] = e.first;
    }
    std::vector<int> queue;
    queue.reserve(nraw);
    for (int b = 0; b < nraw; ++b)
        if (!indeg[b]) queue.push_back(b);
    for (size_t h = 0; h < queue.size(); ++h) {
        const int b = queue[h];
        for (int k = sptr[b]; k < sptr[b + 1]; ++k) {
            const int t = succ[k];
            blev[t] = std::max(blev[t], blev[b] + 1);
            if (--indeg[t] == 0) queue.push_back(t);
        }
    }
    if (static_cast<int>(queue.size()) != nraw) return HFH_OK;   // a cycle
    // non-empty blocks in (level, block) order
    int nlev = 0;
    for (int b = 0; b < nraw; ++b) nlev = std::max(nlev, blev[b] + 1);
    std::vector<int> lcnt(nlev + 1, 0);
    for (int b = 0; b < nraw; ++b)
        if (bcnt[b + 1] > bcnt[b]) ++lcnt[blev[b] + 1];
    for (int l = 0; l < nlev; ++l) lcnt[l + 1] += lcnt[l];
    const int nblk = lcnt[nlev];
    std::vector<int> order(nblk);
    {
        std::vector<int> f(lcnt.begin(), lcnt.end() - 1);
        for (int b = 0; b < nraw; ++b)
            if (bcnt[b + 1] > bcnt[b]) order[f[blev[b]]++] = b;
    }
    W.lslice.assign(lcnt.begin(), lcnt.end());
    const size_t nl64 = 64ULL * std::max(nblk, 1);
    std::vector<int> xptr(nblk + 1, 0), xrow(nl64, -1), xlen(nl64, 0), xlev(nl64, -1), xdpos(nl64, -1),
        xnl(std::max(nblk, 1), 0);
    long long total = 0;
    for (int t = 0; t < nblk; ++t) {
        const int b = order[t];
        int L = 0, nl = 0;
        for (int k = bcnt[b]; k < bcnt[b + 1]; ++k) {
            const int i = brows[k], p = t * 64 + lane[i];
            int cnt = 0;
            for (int j = row[i]; j < row[i + 1]; ++j) {
                if (col[j] == i) xdpos[p] = j;
                cnt += dep(col[j], i);
            }
            xrow[p] = i;
            xlen[p] = cnt;
            xlev[p] = ilev[i];
            L = std::max(L, cnt);
            nl = std::max(nl, ilev[i] + 1);
        }
        xnl[t] = nl;
        xptr[t] = static_cast<int>(total);
        total += 64LL * L;
        HFH_CHECK(total < (1LL << 31), HFH_ERR_UNSUPPORTED, "sweep copy exceeds 2^31 entries");
    }
    xptr[nblk] = static_cast<int>(total);
    std::vector<int> xcol(std::max(total, 1LL), 0), xpos(std::max(total, 1LL), -1);
    for (int t = 0; t < nblk; ++t)
        for (int ln = 0; ln < 64; ++ln) {
            const int p = t * 64 + ln, i = xrow[p];
            if (i < 0) continue;
            int k = 0;
            for (int j = row[i]; j < row[i + 1]; ++j)
                if (dep(col[j], i)) {
                    const long long idx = xptr[t] + 64LL * k + ln;
                    xcol[idx] = bid[col[j]] == bid[i] ? -1 - lane[col[j]] : col[j];
                    xpos[idx] = j;
                    ++k;
                }
        }
    W.nslice = nblk;
    W.nent = total;
    W.kmax = kmax;
    int rc;
    if ((rc = dev_alloc(reinterpret_cast<void **>(&W.ptr), sizeof(int) * xptr.size())) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&W.row), sizeof(int) * xrow.size())) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&W.len), sizeof(int) * xlen.size())) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&W.lev), sizeof(int) * xlev.size())) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&W.bnl), sizeof(int) * xnl.size())) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&W.dpos), sizeof(int) * xdpos.size())) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&W.col), sizeof(int) * xcol.size())) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&W.pos), sizeof(int) * xpos.size())) ||
        (rc = dev_alloc(&W.val, A->elem() * xcol.size())) || (rc = dev_alloc(&W.dval, A->elem() * xdpos.size())))
        return rc;
    HFH_HIP(hipMemcpy(W.ptr, xptr.data(), sizeof(int) * xptr.size(), hipMemcpyHostToDevice));
    HFH_HIP(hipMemcpy(W.row, xrow.data(), sizeof(int) * xrow.size(), hipMemcpyHostToDevice));
    HFH_HIP(hipMemcpy(W.len, xlen.data(), sizeof(int) * xlen.size(), hipMemcpyHostToDevice));
    HFH_HIP(hipMemcpy(W.lev, xlev.data(), sizeof(int) * xlev.size(), hipMemcpyHostToDevice));
    HFH_HIP(hipMemcpy(W.bnl, xnl.data(), sizeof(int) * xnl.size(), hipMemcpyHostToDevice));
    HFH_HIP(hipMemcpy(W.dpos, xdpos.data(), sizeof(int) * xdpos.size(), hipMemcpyHostToDevice));
    HFH_HIP(hipMemcpy(W.col, xcol.data(), sizeof(int) * xcol.size(), hipMemcpyHostToDevice));
    HFH_HIP(hipMemcpy(W.pos, xpos.data(), sizeof(int) * xpos.size(), hipMemcpyHostToDevice));
    return tsweep_fill(A, W);
}

static void free_tsweep(hfh_tsweep &W) {
    void *bufs[] = {W.lslice_d, W.ptr, W.row, W.len, W.col, W.pos, W.dpos, W.val, W.dval, W.lev, W.bnl};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    W = hfh_tsweep();
}

static void free_sweep(hfh_sweep &W) {
    if (W.ptr) (void)hipFree(W.ptr);
    if (W.len) (void)hipFree(W.len);
    if (W.col) (void)hipFree(W.col);
    if (W.val) (void)hipFree(W.val);
    if (W.code) (void)hipFree(W.code);
    if (W.wptr) (void)hipFree(W.wptr);
    if (W.dict) (void)hipFree(W.dict);
    W = hfh_sweep();
}

// Applies the level-scheduled sweeps through a captured graph of every level
// launch (built on first use); the sweeps run on the internal buffers gin /
// gout, with one copy in and one copy out.  Falls back to direct launches if
// the stream cannot be captured.
template <typename T>
static int trsv_graph(hfh_prec *P, const T *in, T *out) {
    hipStream_t s = P->ctx->stream;
    // three launches already: no graph
    if (P->sweep_mode != HFH_SWEEP_LEVELS && P->sf_y && in != out) return trsv_sweeps<T>(P, in, out, nullptr);
    const size_t bytes = sizeof(T) * static_cast<size_t>(P->n);
    if (!P->sweep_exec && !P->graph_failed && graphs_enabled()) {
        int rc = HFH_OK;
        if (!P->gin && (rc = dev_alloc(&P->gin, bytes))) return rc;
        if (!P->gout && (rc = dev_alloc(&P->gout, bytes))) return rc;
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        bool ok = hipStreamIsCapturing(s, &cap) == hipSuccess && cap == hipStreamCaptureStatusNone &&
                  hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
            rc = trsv_sweeps<T>(P, static_cast<const T *>(P->gin), static_cast<T *>(P->gout), nullptr);
            ok = hipStreamEndCapture(s, &P->sweep_graph) == hipSuccess && rc == HFH_OK &&
                 hipGraphInstantiate(&P->sweep_exec, P->sweep_graph, nullptr, nullptr, 0) == hipSuccess;
        }
        if (!ok) {
            (void)hipGetLastError();
            if (P->sweep_exec) (void)hipGraphExecDestroy(P->sweep_exec);
            if (P->sweep_graph) (void)hipGraphDestroy(P->sweep_graph);
            P->sweep_exec = nullptr;
            P->sweep_graph = nullptr;
            P->graph_failed = true;
        }
    }
    if (!P->sweep_exec) return trsv_sweeps<T>(P, in, out, nullptr);
    HFH_HIP(hipMemcpyAsync(P->gin, in, bytes, hipMemcpyDeviceToDevice, s));
    HFH_HIP(hipGraphLaunch(P->sweep_exec, s));
    HFH_HIP(hipMemcpyAsync(out, P->gout, bytes, hipMemcpyDeviceToDevice, s));
    return HFH_OK;
}

int prec_sync(hfh_prec *P) {
    if (!P || (P->kind != HFH_PREC_SGS && P->kind != HFH_PREC_ILU0) || P->tversion == P->A->version) return HFH_OK;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(P->ctx->stream, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return HFH_OK;
    int rc = tsweep_fill(P->A, P->tlo);
    if (!rc) rc = tsweep_fill(P->A, P->tup);
    if (!rc && P->blo.nslice > 0) rc = tsweep_fill(P->A, P->blo);
    if (!rc && P->bup.nslice > 0) rc = tsweep_fill(P->A, P->bup);
    if (!rc) P->tversion = P->A->version;
    return rc;
}

int prec_apply_raw(hfh_prec *P, const void *in, void *out, const int *done, bool mc_first_done, double *rzp,
                   double *rz_out) {
    hipStream_t s = P->ctx->stream;
    const int n = P->n;
    if (n == 0) return HFH_OK;
    int rc = prec_sync(P);
    if (rc) return rc;
    switch (P->kind) {
    case HFH_PREC_NONE:
        HFH_CHECK(!done, HFH_ERR_INVALID, "prec_apply: identity preconditioner inside a driver");
        HFH_HIP(hipMemcpyAsync(out, in, (P->dtype == HFH_F64 ? 8 : 4) * static_cast<size_t>(n), hipMemcpyDeviceToDevice, s));
        return HFH_OK;
    case HFH_PREC_JACOBI: {
        const int nb = std::min((n + kBlock - 1) / kBlock, std::max(P->ctx->num_cu, 1) * 16);
        if (P->dtype == HFH_F64)
            hipLaunchKernelGGL(k_jacobi<double>, dim3(nb), dim3(kBlock), 0, s, n, (const double *)in, (const double *)P->invd, (double *)out, done);
        else
            hipLaunchKernelGGL(k_jacobi<float>, dim3(nb), dim3(kBlock), 0, s, n, (const float *)in, (const float *)P->invd, (float *)out, done);
        return launch_check("k_jacobi");
    }
    case HFH_PREC_MCSGS:
        return P->dtype == HFH_F64
                   ? mcsgs_apply<double>(P, (const double *)in, (double *)out, done, mc_first_done, rzp, rz_out)
                   : mcsgs_apply<float>(P, (const float *)in, (float *)out, done, mc_first_done, rzp, rz_out);
    case HFH_PREC_SGS:
    case HFH_PREC_ILU0:
        if (!done)
            return P->dtype == HFH_F64 ? trsv_graph<double>(P, (const double *)in, (double *)out)
                                       : trsv_graph<float>(P, (const float *)in, (float *)out);
        return P->dtype == HFH_F64 ? trsv_sweeps<double>(P, (const double *)in, (double *)out, done)
                                   : trsv_sweeps<float>(P, (const float *)in, (float *)out, done);
    default:
        set_error("prec_apply: unknown kind %d", P->kind);
        return HFH_ERR_INVALID;
    }
}

}  // namespace hfh

using namespace hfh;

static hfh_prec *new_prec(const hfh_csr *A, int kind) {
    hfh_prec *P = new hfh_prec();
    P->kind = kind;
    P->ctx = A->ctx;
    ++P->ctx->live;
    P->A = A;
    P->n = A->nrow;
    P->dtype = A->dtype;
    return P;
}

extern "C" int hfh_jacobi_create(const hfh_csr *A, hfh_prec **out) {
    HFH_CHECK(A && out && A->has_structure, HFH_ERR_INVALID, "hfh_jacobi_create: bad argument");
    HFH_CHECK(A->nrow == A->ncol, HFH_ERR_INVALID, "hfh_jacobi_create: matrix is not square");
    hfh_prec *P = new_prec(A, HFH_PREC_JACOBI);
    int rc = dev_alloc(&P->invd, A->elem() * static_cast<size_t>(A->nrow));
    if (!rc) {
        hfh_vec tmp;
        tmp.ctx = A->ctx;
        tmp.n = A->nrow;
        tmp.dtype = A->dtype;
        tmp.d = P->invd;
        rc = hfh_csr_invdiag(A, &tmp);
        if (!rc) rc = count_bad_diag(A, "hfh_jacobi_create", nullptr, nullptr, 1.0);
    }
    if (rc) { hfh_prec_destroy(P); return rc; }
    *out = P;
    return HFH_OK;
}

extern "C" int hfh_mcsgs_create(const hfh_csr *A, int ncolors, const int *color_sizes, double omega, hfh_prec **out) {
    HFH_CHECK(A && out && A->has_structure && color_sizes, HFH_ERR_INVALID, "hfh_mcsgs_create: bad argument");
    HFH_CHECK(A->nrow == A->ncol, HFH_ERR_INVALID, "hfh_mcsgs_create: matrix is not square");
    HFH_CHECK(A->sorted, HFH_ERR_INVALID,
              "hfh_mcsgs_create: columns must be ascending in every row (not a diagonal-first matrix)");
    HFH_CHECK(ncolors >= 1 && ncolors <= kMaxColors, HFH_ERR_UNSUPPORTED, "hfh_mcsgs_create: %d colours (1..%d supported)",
              ncolors, kMaxColors);
    HFH_CHECK(omega > 0.0 && omega < 2.0, HFH_ERR_INVALID, "hfh_mcsgs_create: omega %g outside (0, 2)", omega);
    hfh_prec *P = new_prec(A, HFH_PREC_MCSGS);
    P->ncolors = ncolors;
    P->omega = omega;
    P->color_off.assign(ncolors + 1, 0);
    for (int c = 0; c < ncolors; ++c) {
        if (color_sizes[c] < 0) { hfh_prec_destroy(P); hfh::set_error("hfh_mcsgs_create: negative colour size"); return HFH_ERR_INVALID; }
        P->color_off[c + 1] = P->color_off[c] + color_sizes[c];
    }
    if (P->color_off[ncolors] != A->nrow) {
        hfh_prec_destroy(P);
        hfh::set_error("hfh_mcsgs_create: colour sizes sum to %d, matrix has %d rows", P->color_off.back(), A->nrow);
        return HFH_ERR_INVALID;
    }
    int rc = dev_alloc(&P->invd, A->elem() * static_cast<size_t>(A->nrow));
    if (!rc) rc = dev_alloc(&P->dv, A->elem() * static_cast<size_t>(A->nrow));
    if (!rc) rc = count_bad_diag(A, "hfh_mcsgs_create", P->invd, P->dv, 1.0 / (omega * (2.0 - omega)));
    if (!rc && omega != 1.0 && A->nnz > 0) {
        rc = dev_alloc(&P->sval, A->elem() * static_cast<size_t>(A->nnz));
        if (!rc) {
            const int nb = (A->nrow + kBlock - 1) / kBlock;
            if (A->dtype == HFH_F64)
                hipLaunchKernelGGL(k_scale_offdiag<double>, dim3(nb), dim3(kBlock), 0, A->ctx->stream, A->nrow, A->row, A->col,
                                   (const double *)A->val, omega, (double *)P->sval);
            else
                hipLaunchKernelGGL(k_scale_offdiag<float>, dim3(nb), dim3(kBlock), 0, A->ctx->stream, A->nrow, A->row, A->col,
                                   (const float *)A->val, (float)omega, (float *)P->sval);
            rc = launch_check("k_scale_offdiag");
        }
    }
    if (!rc) rc = build_sweep(P, A, true, P->fw);
    if (!rc) rc = build_sweep(P, A, false, P->bw);
    if (rc) { hfh_prec_destroy(P); return rc; }
    *out = P;
    return HFH_OK;
}

static int create_trsv(const hfh_csr *A, int kind, int nlev, const int *lp, const int *lr, int nlu, const int *up,
                       const int *ur, hfh_prec **out, const char *what) {
    HFH_CHECK(A && out && A->has_structure, HFH_ERR_INVALID, "%s: bad argument", what);
    HFH_CHECK(A->nrow == A->ncol, HFH_ERR_INVALID, "%s: matrix is not square", what);
    HFH_CHECK(A->sorted, HFH_ERR_INVALID, "%s: columns must be ascending in every row (not a diagonal-first matrix)",
              what);
    hfh_prec *P = new_prec(A, kind);
    int rc = count_bad_diag(A, what, nullptr, nullptr, 1.0);
    if (!rc) rc = upload_levels(P->lo, A->nrow, nlev, lp, lr, what);
    if (!rc) rc = upload_levels(P->up, A->nrow, nlu, up, ur, what);
    if (!rc) rc = validate_levels(A, P->lo, false, what);
    if (!rc) rc = validate_levels(A, P->up, true, what);
    if (!rc) rc = build_tsweep(A, P->lo, lr, true, P->tlo);
    if (!rc) rc = build_tsweep(A, P->up, ur, false, P->tup);
    int gnx = 0;
    long long gnxy = 0;
    if (!rc && detect_grid(A, gnx, gnxy)) {
        for (int shear = 0; shear < 2 && !rc && P->blo.nslice == 0; ++shear)
            rc = build_bsweep(A, gnx, gnxy, true, shear, P->blo);
        for (int shear = 0; shear < 2 && !rc && P->bup.nslice == 0; ++shear)
            rc = build_bsweep(A, gnx, gnxy, false, shear, P->bup);
        // a block level costs about two row levels (one round trip, then
        // the block's own levels): keep the blocks where they halve the
        // chain at least
        auto keep = [](const hfh_tsweep &B, int nlev) {
            return B.nslice > 0 && 2 * (static_cast<long long>(B.lslice.size()) - 1) <= nlev;
        };
        if (!keep(P->blo, P->lo.nlev)) free_tsweep(P->blo);
        if (!keep(P->bup, P->up.nlev)) free_tsweep(P->bup);
    }
    // sync-free sweeps: forward results and the ticket / failure words
    if (!rc) rc = dev_alloc(&P->sf_y, A->elem() * static_cast<size_t>(std::max(A->nrow, 1)));
    if (!rc) rc = dev_alloc(reinterpret_cast<void **>(&P->sf_ctl), sizeof(int) * kSfCtlInts);
    if (!rc) {
        const hipError_t e = hipMemsetAsync(P->sf_ctl, 0, sizeof(int) * kSfCtlInts, A->ctx->stream);
        if (e != hipSuccess) rc = hip_fail(e, what);
    }
    if (!rc) {
        P->tversion = A->version;
        const hipError_t e = hipStreamSynchronize(A->ctx->stream);
        if (e != hipSuccess) rc = hip_fail(e, what);
    }
    if (rc) { hfh_prec_destroy(P); return rc; }
    *out = P;
    return HFH_OK;
}

extern "C" int hfh_sgs_create(const hfh_csr *A, int nlev, const int *lp, const int *lr, int nlu, const int *up,
                              const int *ur, hfh_prec **out) {
    return create_trsv(A, HFH_PREC_SGS, nlev, lp, lr, nlu, up, ur, out, "hfh_sgs_create");
}

extern "C" int hfh_ilu0_create(const hfh_csr *LU, int nlev, const int *lp, const int *lr, int nlu, const int *up,
                               const int *ur, hfh_prec **out) {
    return create_trsv(LU, HFH_PREC_ILU0, nlev, lp, lr, nlu, up, ur, out, "hfh_ilu0_create");
}

extern "C" int hfh_prec_apply(hfh_prec *P, const hfh_vec *in, hfh_vec *out) {
    HFH_CHECK(P && in && out, HFH_ERR_INVALID, "hfh_prec_apply: null argument");
    HFH_CHECK(in->n == P->n && out->n == P->n && in->dtype == P->dtype && out->dtype == P->dtype, HFH_ERR_INVALID,
              "hfh_prec_apply: vector size/dtype mismatch (n %d, in %d, out %d)", P->n, in->n, out->n);
    HFH_CHECK(in->d != out->d, HFH_ERR_INVALID, "hfh_prec_apply: in and out must not alias");
    return prec_apply_raw(P, in->d, out->d);
}

extern "C" int hfh_prec_sweep_blocks(const hfh_prec *P, int *fw_blocks, int *fw_levels, int *bw_blocks,
                                     int *bw_levels) {
    HFH_CHECK(P && fw_blocks && fw_levels && bw_blocks && bw_levels, HFH_ERR_INVALID,
              "hfh_prec_sweep_blocks: null argument");
    auto lv = [](const hfh_tsweep &B) { return B.nslice > 0 ? static_cast<int>(B.lslice.size()) - 1 : 0; };
    *fw_blocks = P->blo.nslice;
    *fw_levels = lv(P->blo);
    *bw_blocks = P->bup.nslice;
    *bw_levels = lv(P->bup);
    return HFH_OK;
}

extern "C" int hfh_prec_set_sweep(hfh_prec *P, int mode, int waves_per_cu) {
    HFH_CHECK(P, HFH_ERR_INVALID, "hfh_prec_set_sweep: null preconditioner");
    HFH_CHECK(mode == HFH_SWEEP_SYNCFREE || mode == HFH_SWEEP_LEVELS || mode == HFH_SWEEP_SLICES, HFH_ERR_INVALID,
              "hfh_prec_set_sweep: mode %d is not HFH_SWEEP_SYNCFREE, HFH_SWEEP_LEVELS or HFH_SWEEP_SLICES", mode);
    HFH_CHECK(waves_per_cu >= 0 && waves_per_cu <= 32, HFH_ERR_INVALID,
              "hfh_prec_set_sweep: waves_per_cu %d not 0 (default) or in [1, 32]", waves_per_cu);
    P->sweep_mode = mode;
    P->sf_waves = waves_per_cu;
    return HFH_OK;
}

int hfh::sweep_failed(hfh_prec *P) {
    if (!P || !P->sf_ctl) return HFH_OK;
    hipStream_t s = P->ctx->stream;
    int *h = reinterpret_cast<int *>(P->ctx->host_res);
    HFH_HIP(hipMemcpyAsync(h, P->sf_ctl + 2, sizeof(int), hipMemcpyDeviceToHost, s));
    HFH_HIP(hipStreamSynchronize(s));
    if (*h == 0) return HFH_OK;
    HFH_HIP(hipMemcpyAsync(h, P->sf_ctl + 3, sizeof(int), hipMemcpyDeviceToHost, s));
    HFH_HIP(hipStreamSynchronize(s));
    const int most = *h;
    HFH_HIP(hipMemsetAsync(P->sf_ctl + 2, 0, 2 * sizeof(int), s));
    set_error("sync-free triangular sweep: a row waited longer than its bound for an earlier row's result "
              "(%d polls; results are NaN)", most);
    return HFH_ERR_HIP;
}

extern "C" int hfh_prec_sweep_status(hfh_prec *P) {
    HFH_CHECK(P, HFH_ERR_INVALID, "hfh_prec_sweep_status: null preconditioner");
    return sweep_failed(P);
}

extern "C" int hfh_prec_kind(const hfh_prec *P, int *kind) {
    HFH_CHECK(P && kind, HFH_ERR_INVALID, "hfh_prec_kind: null argument");
    *kind = P->kind;
    return HFH_OK;
}

extern "C" int hfh_prec_destroy(hfh_prec *P) {
    if (!P) return HFH_OK;
    if (P->ctx && P->ctx->stream) (void)hipStreamSynchronize(P->ctx->stream);
    if (P->ctx) --P->ctx->live;
    if (P->invd) (void)hipFree(P->invd);
    if (P->dv) (void)hipFree(P->dv);
    if (P->sval) (void)hipFree(P->sval);
    if (P->lo.rows) (void)hipFree(P->lo.rows);
    if (P->up.rows) (void)hipFree(P->up.rows);
    if (P->lo.ptr_d) (void)hipFree(P->lo.ptr_d);
    if (P->up.ptr_d) (void)hipFree(P->up.ptr_d);
    free_sweep(P->fw);
    free_sweep(P->bw);
    free_tsweep(P->tlo);
    free_tsweep(P->tup);
    free_tsweep(P->blo);
    free_tsweep(P->bup);
    if (P->sweep_exec) (void)hipGraphExecDestroy(P->sweep_exec);
    if (P->sweep_graph) (void)hipGraphDestroy(P->sweep_graph);
    if (P->gin) (void)hipFree(P->gin);
    if (P->gout) (void)hipFree(P->gout);
    if (P->sf_y) (void)hipFree(P->sf_y);
    if (P->sf_ctl) (void)hipFree(P->sf_ctl);
    delete P;
    return HFH_OK;
}

// ILU(0) factorisation on the device, in place, level by level along the
// lower level schedule (the rows of one level do not depend on each other).
extern "C" int hfh_csr_ilu0(hfh_csr *A, int nlev, const int *level_ptr, const int *level_rows, int *missing) {
    HFH_CHECK(A && A->has_structure && A->nrow == A->ncol, HFH_ERR_INVALID, "hfh_csr_ilu0: bad matrix");
    HFH_CHECK(A->sorted, HFH_ERR_INVALID,
              "hfh_csr_ilu0: columns must be ascending in every row (not a diagonal-first matrix)");
    hfh_levels L;
    int rc = upload_levels(L, A->nrow, nlev, level_ptr, level_rows, "hfh_csr_ilu0");
    if (!rc) rc = validate_levels(A, L, false, "hfh_csr_ilu0");
    if (rc) {
        if (L.rows) (void)hipFree(L.rows);
        if (L.ptr_d) (void)hipFree(L.ptr_d);
        return rc;
    }
    // lower entries whose row k has no stored diagonal are skipped, as on the host
    int miss = 0;
    {
        std::vector<char> hasd(A->nrow, 0);
        for (int i = 0; i < A->nrow; ++i)
            for (int j = A->h_row[i]; j < A->h_row[i + 1]; ++j)
                if (A->h_col[j] == i) hasd[i] = 1;
        for (int i = 0; i < A->nrow; ++i)
            for (int j = A->h_row[i]; j < A->h_row[i + 1] && A->h_col[j] < i; ++j)
                if (!hasd[A->h_col[j]]) ++miss;
    }
    hipStream_t s = A->ctx->stream;
    for (const auto &run : L.runs) {
        if (run.second - run.first > 1) {
            if (A->dtype == HFH_F64)
                hipLaunchKernelGGL(k_ilu0_run<double>, dim3(1), dim3(kMergeBlock), 0, s, L.rows, L.ptr_d, run.first,
                                   run.second, A->row, A->col, (double *)A->val, A->diag);
            else
                hipLaunchKernelGGL(k_ilu0_run<float>, dim3(1), dim3(kMergeBlock), 0, s, L.rows, L.ptr_d, run.first,
                                   run.second, A->row, A->col, (float *)A->val, A->diag);
            continue;
        }
        const int a = L.ptr[run.first], cnt = L.ptr[run.first + 1] - a;
        if (cnt == 0) continue;
        const dim3 g((cnt + kBlock - 1) / kBlock);
        if (A->dtype == HFH_F64)
            hipLaunchKernelGGL(k_ilu0_level<double>, g, dim3(kBlock), 0, s, L.rows + a, cnt, A->row, A->col,
                               (double *)A->val, A->diag);
        else
            hipLaunchKernelGGL(k_ilu0_level<float>, g, dim3(kBlock), 0, s, L.rows + a, cnt, A->row, A->col,
                               (float *)A->val, A->diag);
    }
    rc = launch_check("k_ilu0_level");
    (void)hipStreamSynchronize(s);
    (void)hipFree(L.rows);
    (void)hipFree(L.ptr_d);
    if (rc) return rc;
    // the SELL copy holds the old values
    ++A->version;
    A->sell_dirty = true;
    if ((rc = spmv_sell_fill(A))) return rc;
    if (missing) *missing = miss;
    return HFH_OK;
}
