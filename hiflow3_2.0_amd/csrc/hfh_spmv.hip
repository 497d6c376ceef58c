// hfh_spmv.hip -- CSR storage and the streaming CSR SpMV for gfx950.
//
// SpMV (lMatrix::VectorMult, CPUsimple_csr_lmatrix.cc:75-105) is HBM-bound:
// per call it must stream val (8 B/nnz) + col (4 B/nnz) + row (4 B/row) and
// write y (8 B/row); x is gathered through L2 / Infinity Cache.
//
// HFH_SPMV_STREAM: the rows are partitioned into row blocks of at most `tile`
// nonzeros.  A workgroup streams a block's val/col with coalesced,
// non-temporal loads (so the streams do not evict the x window from L2),
// forms the products val[j]*x[col[j]] in LDS, and then one lane per row sums
// its products in ascending j.  The products are rounded before the sum
// exactly like the reference's `res += val[j]*x[col[j]]`, so rows that fit in
// one block agree bit for bit with the CPU path.  Rows longer than a block
// get a block of their own and a tree sum.
// HFH_SPMV_VECTOR: GL lanes per row, shuffle-tree row sums (not bitwise).
// Both are persistent over an XCD-contiguous split: each XCD streams one
// contiguous eighth of the matrix, so its L2 holds one x window.
#include "hfh_internal.h"

#include <algorithm>
#include <cstdint>
#include <thread>

namespace hfh {

// Work split shared by both SpMV kernels: with G >= 8 workgroups (G a
// multiple of 8, consecutive workgroup ids land on consecutive XCDs) XCD x
// owns the contiguous item range [x*items/8, (x+1)*items/8) and its G/8
// workgroups stride through it; below 8 workgroups the items are strided.
struct Split {
    int lo, hi, first, step;
};

__device__ __forceinline__ Split xcd_split(int items, int per) {
    const int G = gridDim.x, g = blockIdx.x;
    Split sp;
    if (G >= kNumXcd) {
        const int x = g % kNumXcd, slot = g / kNumXcd, nslot = G / kNumXcd;
        sp.lo = static_cast<int>(static_cast<long long>(items) * x / kNumXcd);
        sp.hi = static_cast<int>(static_cast<long long>(items) * (x + 1) / kNumXcd);
        sp.first = sp.lo + slot * per;
        sp.step = nslot * per;
    } else {
        sp.lo = 0;
        sp.hi = items;
        sp.first = g * per;
        sp.step = G * per;
    }
    return sp;
}

// Row-block streaming kernel.  Each workgroup walks its row blocks; a block's
// val/col stream is loaded coalesced (non-temporal), multiplied by the
// gathered x into LDS, and each row is then summed by one lane in ascending j.
template <typename T, bool ADD, bool DOT, int TILE>
__global__ __launch_bounds__(kBlock) void k_spmv_stream(
    const int *__restrict__ row, const int *__restrict__ col, const T *__restrict__ val,
    const T *__restrict__ x, T *__restrict__ y, const int *__restrict__ rblk, int nblk,
    const T *__restrict__ w, double *part, const int *__restrict__ done) {
    if (done && *done) return;
    __shared__ T prod[TILE];
    __shared__ T sht[kBlock / 64];
    __shared__ double shd[kBlock / 64];
    const Split sp = xcd_split(nblk, 1);
    double dsum = 0;
    for (int b = sp.first; b < sp.hi; b += sp.step) {
        const int r0 = rblk[b], r1 = rblk[b + 1];
        const int j0 = row[r0], j1 = row[r1];
        const int cnt = j1 - j0;
        if (cnt <= TILE) {
            // unconditional loads at clamped positions, so every load of the
            // tile is in flight before the first gather result is needed; a
            // block of empty rows (cnt == 0, uniform over the workgroup) loads
            // nothing: its clamp target col[j0] may lie past the last entry
            if (cnt > 0) {
                constexpr int K = TILE / kBlock;
                int c[K];
                T v[K], xv[K];
                const int last = cnt - 1;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int e = min(static_cast<int>(threadIdx.x) + k * kBlock, last);
                    c[k] = __builtin_nontemporal_load(&col[j0 + e]);
                    v[k] = __builtin_nontemporal_load(&val[j0 + e]);
                }
#pragma unroll
                for (int k = 0; k < K; ++k) xv[k] = x[c[k]];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int e = threadIdx.x + k * kBlock;
                    if (e < cnt) prod[e] = v[k] * xv[k];
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
            __syncthreads();
        } else {
            // one long row (r1 == r0 + 1): strided partial sums, fixed tree
            T s = 0;
            for (int j = j0 + threadIdx.x; j < j1; j += kBlock) s += val[j] * x[col[j]];
            s = block_sum(s, sht);
            if (threadIdx.x == 0) {
                if (ADD) s = y[r0] + s;
                y[r0] = s;
                if (DOT) dsum += static_cast<double>(s) * static_cast<double>(w[r0]);
            }
            __syncthreads();
        }
    }
    if (DOT) {
        const double v[1] = {block_sum(dsum, shd)};
        store_partials<1>(v, part, 0, blockIdx.x);
    }
}

// Wave-tile kernel: each wavefront owns its tiles (at most WT nonzeros and
// 63 rows, so the 64 lanes hold the row pointers) and a private LDS slice, so waves never wait for each other.  The
// tile's row pointers come in with one coalesced load; the val/col loads of
// the wave's next tile are issued before the current tile's row sums, so a
// wave keeps a tile of HBM reads in flight while it reduces.  Row sums are
// sequential in ascending j over the rounded products (bit-exact).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T, bool ADD, bool DOT, int WT>
__global__ __launch_bounds__(kBlock) void k_spmv_wave(
    const int *__restrict__ row, const int *__restrict__ col, const T *__restrict__ val,
    const T *__restrict__ x, T *__restrict__ y, const int *__restrict__ rblk, int nblk,
    const T *__restrict__ w, double *part, const int *__restrict__ done) {
    if (done && *done) return;
    constexpr int WPB = kBlock / 64, PER = WT / 64;
    __shared__ T prod_all[WPB][WT];
    __shared__ double shd[kBlock / 64];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    T *prod = prod_all[wave];
    const Split sp = xcd_split(nblk, WPB);
    double dsum = 0;
    int b = sp.first + wave;
    // registers of the tile being loaded
    T v[PER];
    int c[PER];
    int j0 = 0, cnt = 0, r0 = 0, r1 = 0;
    auto issue = [&](int bb) {
        r0 = rblk[bb];
        r1 = rblk[bb + 1];
        j0 = row[r0];
        cnt = row[r1] - j0;
        if (cnt > WT) return;  // a long row is streamed when it is processed
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int e = lane + 64 * k;
            if (e < cnt) {
                c[k] = __builtin_nontemporal_load(&col[j0 + e]);
                v[k] = __builtin_nontemporal_load(&val[j0 + e]);
            }
        }
    };
    if (b < sp.hi) issue(b);
    while (b < sp.hi) {
        const int cr0 = r0, cr1 = r1, cj0 = j0, ccnt = cnt;
        const int bn = b + sp.step;
        if (ccnt > WT) {
            // one long row (cr1 == cr0 + 1): lane-strided sums, fixed shuffle tree
            T s = 0;
            for (int j = cj0 + lane; j < cj0 + ccnt; j += 64) s += val[j] * x[col[j]];
            s = wave_sum(s);
            if (lane == 0) {
                if (ADD) s = y[cr0] + s;
                y[cr0] = s;
                if (DOT) dsum += static_cast<double>(s) * static_cast<double>(w[cr0]);
            }
            if (bn < sp.hi) issue(bn);
            b = bn;
            continue;
        }
        // the row pointers of this tile (<= 63 rows + 1)
        const int rp = (cr0 + lane <= cr1) ? row[cr0 + lane] - cj0 : 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int e = lane + 64 * k;
            if (e < ccnt) prod[e] = v[k] * x[c[k]];
        }
        if (bn < sp.hi) issue(bn);
        wave_sync();
        const int nr = cr1 - cr0;
        const int a = rp;
        const int e = __shfl_down(rp, 1, 64);
        if (lane < nr) {
            T s = 0;
            for (int j = a; j < e; ++j) s += prod[j];
            const int r = cr0 + lane;
            if (ADD) s = y[r] + s;
            y[r] = s;
            if (DOT) dsum += static_cast<double>(s) * static_cast<double>(w[r]);
        }
        wave_sync();
        b = bn;
    }
    if (DOT) {
        __syncthreads();
        const double vv[1] = {block_sum(dsum, shd)};
        store_partials<1>(vv, part, 0, blockIdx.x);
    }
}

// SELL-64 kernel: the matrix is held a second time in slices of 64 rows (one
// wavefront), each slice column-major and padded to its longest row.  Lane =
// row: the val/col streams of a slice are contiguous for the whole wave, and
// the x gathers of one column slot come from neighbouring rows, hence from
// neighbouring addresses.  No LDS, no barriers.  Each row is summed in
// ascending column order with separately rounded products and sums
// (mul_rn / add_rn are compiled without contraction), so rows agree bit for
// bit with the CPU path.

template <typename T, bool ADD, bool DOT, int U>
__global__ __launch_bounds__(kBlock) void k_spmv_sell(
    int nrow, int nslice, const int *__restrict__ sptr, const int *__restrict__ rlen, const int *__restrict__ scol,
    const T *__restrict__ sval, const T *__restrict__ x, T *__restrict__ y, const T *__restrict__ w, double *part,
    const int *__restrict__ done) {
    if (done && *done) return;
    __shared__ double shd[kBlock / 64];
    constexpr int WPB = kBlock / 64;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    double dsum = 0;
    // consecutive workgroups take consecutive slices (measured faster on
    // gfx950 than one contiguous range per XCD: the streams of all XCDs then
    // stay within one moving window of HBM)
    for (int s = blockIdx.x * WPB + wave; s < nslice; s += gridDim.x * WPB) {
        const int base = sptr[s];
        const int L = (sptr[s + 1] - base) >> 6;
        const int r = s * 64 + lane;
        const int len = r < nrow ? rlen[r] : 0;
        T acc = 0;
        // The last step re-reads column slot L-1 for u beyond the slice width,
        // so every step issues its U loads together; those slots (k+u >= L >=
        // len) are masked out of the sum.
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
                if (k + u < len) acc = add_rn(acc, mul_rn(v[u], xv[u]));
        }
        if (r < nrow) {
            if (ADD) acc = y[r] + acc;
            y[r] = acc;
            if (DOT) dsum += static_cast<double>(acc) * static_cast<double>(w[r]);
        }
    }
    if (DOT) {
        __syncthreads();
        const double vv[1] = {block_sum(dsum, shd)};
        store_partials<1>(vv, part, 0, blockIdx.x);
    }
}

template <typename T> struct Pair;
template <> struct Pair<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct Pair<float> { typedef float type __attribute__((ext_vector_type(2))); };
typedef int int_pair __attribute__((ext_vector_type(2)));

// SELL-64 with slot pairs: slots 2q, 2q+1 of a lane at base + q*128 + 2*lane,
// so each lane loads its two values (and columns) with one 16-byte (8-byte)
// load.  A slice of odd width L keeps its last slot single, at
// base + (L/2)*128 + lane, so no slice is padded beyond its longest row.
// Row sums as k_spmv_sell (ascending slots, uncontracted).
template <typename T, bool ADD, bool DOT, int U>
__global__ __launch_bounds__(kBlock) void k_spmv_sellp(
    int nrow, int nslice, const int *__restrict__ sptr, const int *__restrict__ rlen, const int *__restrict__ scol,
    const T *__restrict__ sval, const T *__restrict__ x, T *__restrict__ y, const T *__restrict__ w, double *part,
    const int *__restrict__ done) {
    typedef typename Pair<T>::type T2;
    if (done && *done) return;
    __shared__ double shd[kBlock / 64];
    constexpr int WPB = kBlock / 64;
    // the wave index in a scalar register: slice quantities are then scalar
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    double dsum = 0;
    for (int s = blockIdx.x * WPB + wave; s < nslice; s += gridDim.x * WPB) {
        const int base = sptr[s], L = (sptr[s + 1] - base) >> 6, Q = L >> 1;   // Q slot pairs
        const int r = s * 64 + lane;
        const int len = r < nrow ? rlen[r] : 0;
        T acc = 0;
        for (int q = 0; q < Q; q += U) {
            int_pair c[U];
            T2 v[U];
            T xa[U], xb[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int idx = base + min(q + u, Q - 1) * 128 + 2 * lane;
                c[u] = __builtin_nontemporal_load(reinterpret_cast<const int_pair *>(&scol[idx]));
                v[u] = __builtin_nontemporal_load(reinterpret_cast<const T2 *>(&sval[idx]));
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                xa[u] = x[c[u].x];
                xb[u] = x[c[u].y];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                // a step past the last pair re-loaded pair Q-1: masked out
                // (2Q may be below len when the odd last slot is single)
                const int k = 2 * (q + u);
                const bool pair = q + u < Q;
                if (pair && k < len) acc = add_rn(acc, mul_rn(static_cast<T>(v[u].x), xa[u]));
                if (pair && k + 1 < len) acc = add_rn(acc, mul_rn(static_cast<T>(v[u].y), xb[u]));
            }
        }
        if (L & 1) {   // wave-uniform
            const int idx = base + Q * 128 + lane;
            const int c = __builtin_nontemporal_load(&scol[idx]);
            const T v = __builtin_nontemporal_load(&sval[idx]);
            const T xv = x[c];
            if (2 * Q < len) acc = add_rn(acc, mul_rn(v, xv));
        }
        if (r < nrow) {
            if (ADD) acc = y[r] + acc;
            y[r] = acc;
            if (DOT) dsum += static_cast<double>(acc) * static_cast<double>(w[r]);
        }
    }
    if (DOT) {
        __syncthreads();
        const double vv[1] = {block_sum(dsum, shd)};
        store_partials<1>(vv, part, 0, blockIdx.x);
    }
}

// SELL pair layout with x staged in LDS.  A wave takes a slice, loads the
// slice's chunks of x (CH elements each, listed ascending at cid[cptr[s]..])
// into its own LDS image (img elements per wave, dynamic shared memory sized
// to the matrix's largest image; IMG >= img bounds the staging registers)
// with 16-byte loads (one piece of 16 bytes per lane and step), and reads
// x_col from the image at the slot's 16-bit code (chunk position * CH +
// col % CH).  The
// first slot pair and the single slot of an odd width are loaded before the
// staging, the next pair while the current one is summed (clamped to the
// last pair, so every load stays in the slice).  Row sums as k_spmv_sellp
// (ascending slots, uncontracted), so rows agree bit for bit with the CPU
// loop.  JAC 1 / 2: a Jacobi step (normal / damped) -- entries whose code is
// the row's diagonal code are skipped and xn_r = jacobi_update(...) is
// stored instead of the row sum.  xal: x is 16-byte aligned.
template <typename T, bool ADD, bool DOT, int JAC, int CH, int IMG>
__global__ __launch_bounds__(kBlock) void k_spmv_sellx(
    int nrow, int ncol, int nslice, const int *__restrict__ sptr, const int *__restrict__ rlen,
    const unsigned short *__restrict__ code, const T *__restrict__ sval, const int *__restrict__ cptr,
    const int *__restrict__ cid, int img, int xal, const T *__restrict__ x, T *__restrict__ y,
    const T *__restrict__ w, double *part, const int *__restrict__ done, const unsigned short *__restrict__ dcode,
    const T *__restrict__ b, const T *__restrict__ invd, T jw, T jomw) {
    typedef typename Pair<T>::type T2;
    typedef typename V16<T>::type TV;
    constexpr int WV = V16<T>::W;              // elements of a 16-byte piece
    constexpr int PPC = CH / WV;               // pieces of a chunk
    constexpr int NP = (IMG / WV + 63) / 64;   // pieces per lane
    static_assert(CH % WV == 0 && IMG % CH == 0, "chunk of whole pieces, image of whole chunks");
    constexpr int WPB = kBlock / 64;
    if (done && *done) return;
    extern __shared__ __align__(16) unsigned char xs_raw[];   // WPB images of img elements
    __shared__ double shd[WPB];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    T *xs = reinterpret_cast<T *>(xs_raw) + wave * img;
    const unsigned *code2 = reinterpret_cast<const unsigned *>(code);
    double dsum = 0;
    for (int s = blockIdx.x * WPB + wave; s < nslice; s += gridDim.x * WPB) {
        const int base = sptr[s], L = (sptr[s + 1] - base) >> 6, Q = L >> 1, qmax = Q > 0 ? Q - 1 : 0;
        const int c0 = cptr[s], npc = (cptr[s + 1] - c0) * PPC;
        const int r = s * 64 + lane;
        const int len = r < nrow ? rlen[r] : 0;
        const unsigned dc = JAC && r < nrow ? dcode[r] : 0xffffu;
        unsigned cw = 0, ct = 0;
        T2 v = T2(0);
        T vt = 0;
        if (Q > 0) {   // wave-uniform
            cw = __builtin_nontemporal_load(&code2[(base >> 1) + lane]);
            v = __builtin_nontemporal_load(reinterpret_cast<const T2 *>(&sval[base + 2 * lane]));
        }
        if (L & 1) {   // wave-uniform
            const int idx = base + Q * 128 + lane;
            ct = __builtin_nontemporal_load(&code[idx]);
            vt = __builtin_nontemporal_load(&sval[idx]);
        }
        int id[NP];
        TV xv[NP];
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const int p = lane + 64 * i;
            id[i] = p < npc ? cid[c0 + p / PPC] : 0;
        }
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const int p = lane + 64 * i;
            const int xi = id[i] * CH + (p % PPC) * WV;
            xv[i] = TV(0);
            if (p < npc && xal && xi + WV <= ncol) xv[i] = *reinterpret_cast<const TV *>(&x[xi]);
        }
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const int p = lane + 64 * i;
            if (p < npc) *reinterpret_cast<TV *>(&xs[p * WV]) = xv[i];
        }
        // pieces read element by element: all of them when x is not 16-byte
        // aligned, else only the one holding x's last entries (a separate
        // loop keeps the 16-byte path above free of its registers)
#pragma unroll 1
        for (int i = 0; i < NP; ++i) {
            const int p = lane + 64 * i;
            const int xi = id[i] * CH + (p % PPC) * WV;
            if (p < npc && !(xal && xi + WV <= ncol))
                for (int t = 0; t < WV; ++t)
                    if (xi + t < ncol) xs[p * WV + t] = x[xi + t];
        }
        // the image is private to this wave, whose LDS accesses execute in
        // order: only the compiler has to keep them in program order
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
        T acc = 0;
        for (int q = 0; q < Q; ++q) {
            const unsigned cq = cw;
            const T2 vq = v;
            const int qn = min(q + 1, qmax);
            cw = __builtin_nontemporal_load(&code2[(base >> 1) + qn * 64 + lane]);
            v = __builtin_nontemporal_load(reinterpret_cast<const T2 *>(&sval[base + qn * 128 + 2 * lane]));
            const unsigned ka = cq & 0xffffu, kb = cq >> 16;
            const T xa = xs[ka], xb = xs[kb];
            const int k = 2 * q;
            if (k < len && (!JAC || ka != dc)) acc = add_rn(acc, mul_rn(static_cast<T>(vq.x), xa));
            if (k + 1 < len && (!JAC || kb != dc)) acc = add_rn(acc, mul_rn(static_cast<T>(vq.y), xb));
        }
        if (L & 1) {
            const T xt = xs[ct];
            if (2 * Q < len && (!JAC || ct != dc)) acc = add_rn(acc, mul_rn(vt, xt));
        }
        if (r < nrow) {
            if (JAC) {
                y[r] = jacobi_update<T, JAC == 2>(acc, x[r], b[r], invd[r], jw, jomw);
            } else {
                if (ADD) acc = y[r] + acc;
                y[r] = acc;
                if (DOT) dsum += static_cast<double>(acc) * static_cast<double>(w[r]);
            }
        }
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
    }
    if (DOT) {
        __syncthreads();
        const double vv[1] = {block_sum(dsum, shd)};
        store_partials<1>(vv, part, 0, blockIdx.x);
    }
}

// FP32 rows are light: waves that take several slices each (8 workgroups per
// CU) beat one slice per wave (FV 192^3: 60.6 us against 70.1 us); FP64
// keeps one slice per wave (103.7 us against 115.8 us at 8 workgroups per CU).
template <typename T> constexpr bool kDictCappedGrid = false;
template <> constexpr bool kDictCappedGrid<float> = true;

// SELL-64 with dictionary-coded columns: column = row + dict[code].  Each
// slice keeps its width L of value slots; code word w of a lane holds the
// codes of slots 4w..4w+3 (byte t = slot 4w+t), ceil(L / 4) words per lane,
// and the slots of the last word past L are neither stored nor loaded.  A
// row's slots past its length hold kPadCode, so no row length is loaded
// (its gather reads x[row], in range, and is not summed).  Row sums as
// k_spmv_sell.
// JAC 1 / 2: a Jacobi step (normal / damped) -- entries with the code of
// offset 0 (code0, the diagonal) are skipped and xn_r = jacobi_update(...)
// is stored instead of the row sum.
template <typename T, bool ADD, bool DOT, int UW, int JAC>
__global__ __launch_bounds__(kBlock) void k_spmv_selld(
    int nrow, int ncol, int nslice, const int *__restrict__ sptr, const int *__restrict__ wptr,
    const unsigned *__restrict__ codes, const int *__restrict__ gdict, int ndict,
    const T *__restrict__ sval, const T *__restrict__ x, T *__restrict__ y, const T *__restrict__ w, double *part,
    const int *__restrict__ done, unsigned code0, const T *__restrict__ b, const T *__restrict__ invd, T jw,
    T jomw) {
    if (done && *done) return;
    __shared__ int dict[256];
    __shared__ double shd[kBlock / 64];
    for (int i = threadIdx.x; i < 256; i += kBlock) dict[i] = i < ndict ? gdict[i] : 0;   // kPadCode: offset 0
    __syncthreads();
    constexpr int WPB = kBlock / 64;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    double dsum = 0;
    if constexpr (!kDictCappedGrid<T>) {   // one slice per wave (FP64): loads in place
        for (int s = blockIdx.x * WPB + wave; s < nslice; s += gridDim.x * WPB) {
            const int vbase = sptr[s], L = (sptr[s + 1] - vbase) >> 6, W = (L + 3) >> 2;   // slots, code words
            const int cbase = wptr[s];
            const int r = s * 64 + lane;
            T acc = 0;
            for (int k = 0; k < W; k += UW) {
                unsigned cw[UW];
                T v[UW][4], xv[UW][4];
#pragma unroll
                for (int u = 0; u < UW; ++u) {
                    const int ww = min(k + u, W - 1);
                    cw[u] = __builtin_nontemporal_load(&codes[cbase + ww * 64 + lane]);
#pragma unroll
                    for (int t = 0; t < 4; ++t)   // slots past the slice width: not stored (wave-uniform test)
                        v[u][t] = 4 * ww + t < L ? __builtin_nontemporal_load(&sval[vbase + (4 * ww + t) * 64 + lane])
                                                 : static_cast<T>(0);
                }
#pragma unroll
                for (int u = 0; u < UW; ++u)
#pragma unroll
                    for (int t = 0; t < 4; ++t) xv[u][t] = x[min(r + dict[(cw[u] >> (8 * t)) & 255u], ncol - 1)];
#pragma unroll
                for (int u = 0; u < UW; ++u)
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        // words past W repeat the last one (the clamped load above)
                        const unsigned cd = (cw[u] >> (8 * t)) & 255u;
                        if (k + u < W && cd != kPadCode && (!JAC || cd != code0))
                            acc = add_rn(acc, mul_rn(v[u][t], xv[u][t]));
                    }
            }
            if (r < nrow) {
                if (JAC) {
                    y[r] = jacobi_update<T, JAC == 2>(acc, x[r], b[r], invd[r], jw, jomw);
                } else {
                    if (ADD) acc = y[r] + acc;
                    y[r] = acc;
                    if (DOT) dsum += static_cast<double>(acc) * static_cast<double>(w[r]);
                }
            }
        }
    } else {
        // Loads of step k (words k..k+UW-1) of the slice at value base vb, code
        // base cb, width L slots / W words; words past W repeat the last one.
        auto load_step = [&](int vb, int cb, int L, int W, int k, unsigned (&cw)[UW], T (&v)[UW][4]) {
#pragma unroll
            for (int u = 0; u < UW; ++u) {
                const int ww = min(k + u, W - 1);
                cw[u] = __builtin_nontemporal_load(&codes[cb + ww * 64 + lane]);
#pragma unroll
                for (int t = 0; t < 4; ++t)   // slots past the slice width: not stored (wave-uniform test)
                    v[u][t] = 4 * ww + t < L ? __builtin_nontemporal_load(&sval[vb + (4 * ww + t) * 64 + lane])
                                             : static_cast<T>(0);
            }
        };
        const int stride = gridDim.x * WPB;
        int s = blockIdx.x * WPB + wave;
        // On the capped FP32 grid a wave walks several slices: the header and
        // first step of its next slice are loaded before the row sums of its
        // current one, so their latency overlaps the current gathers.
        int nvb = 0, ncb = 0, nL = 0, nW = 0;
        unsigned ncw[UW];
        T nv[UW][4];
        auto fetch = [&](int sn) {
            nvb = sptr[sn];
            nL = (sptr[sn + 1] - nvb) >> 6;
            nW = (nL + 3) >> 2;
            ncb = wptr[sn];
            if (nW > 0) load_step(nvb, ncb, nL, nW, 0, ncw, nv);
        };
        if (s < nslice) fetch(s);
        for (; s < nslice; s += stride) {
            const int vbase = nvb, L = nL, W = nW, cbase = ncb;   // slots, code words
            const int r = s * 64 + lane;
            T acc = 0;
            for (int k = 0; k < W; k += UW) {
                unsigned cw[UW];
                T v[UW][4], xv[UW][4];
                if (k == 0) {
#pragma unroll
                    for (int u = 0; u < UW; ++u) {
                        cw[u] = ncw[u];
#pragma unroll
                        for (int t = 0; t < 4; ++t) v[u][t] = nv[u][t];
                    }
                } else {
                    load_step(vbase, cbase, L, W, k, cw, v);
                }
#pragma unroll
                for (int u = 0; u < UW; ++u)
#pragma unroll
                    for (int t = 0; t < 4; ++t) xv[u][t] = x[min(r + dict[(cw[u] >> (8 * t)) & 255u], ncol - 1)];
                if (k + UW >= W && s + stride < nslice) fetch(s + stride);   // after this step's gathers
#pragma unroll
                for (int u = 0; u < UW; ++u)
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        // words past W repeat the last one (the clamped load above)
                        const unsigned cd = (cw[u] >> (8 * t)) & 255u;
                        if (k + u < W && cd != kPadCode && (!JAC || cd != code0))
                            acc = add_rn(acc, mul_rn(v[u][t], xv[u][t]));
                    }
            }
            if (W == 0 && s + stride < nslice) fetch(s + stride);   // an empty slice ran no step
            if (r < nrow) {
                if (JAC) {
                    y[r] = jacobi_update<T, JAC == 2>(acc, x[r], b[r], invd[r], jw, jomw);
                } else {
                    if (ADD) acc = y[r] + acc;
                    y[r] = acc;
                    if (DOT) dsum += static_cast<double>(acc) * static_cast<double>(w[r]);
                }
            }
        }
    }
    if (DOT) {
        __syncthreads();
        const double vv[1] = {block_sum(dsum, shd)};
        store_partials<1>(vv, part, 0, blockIdx.x);
    }
}

// One thread per SELL row slot (nslice*64 of them, including the padding rows
// of the last slice).  Padding entries get value 0 and a valid column
// (min(row, ncol-1)), so the kernel can load them without a bounds test.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_csr_to_sell(int nrow, int ncol, int nslice, const int *__restrict__ row,
                                                        const int *__restrict__ col, const T *__restrict__ val,
                                                        const int *__restrict__ sptr, int *__restrict__ scol,
                                                        T *__restrict__ sval, int pairs) {
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= nslice * 64) return;
    const int s = r >> 6, lane = r & 63;
    const int base = sptr[s], L = (sptr[s + 1] - base) >> 6;
    const int j0 = r < nrow ? row[r] : 0, len = r < nrow ? row[r + 1] - j0 : 0;
    const int pad = min(r, ncol - 1);
    const int P = pairs ? (L & ~1) : 0;   // slots held in pairs; an odd last slot is single
    for (int k = 0; k < L; ++k) {
        const int idx = k < P ? base + (k >> 1) * 128 + 2 * lane + (k & 1) : base + k * 64 + lane;
        if (scol) scol[idx] = k < len ? col[j0 + k] : pad;   // null for the dictionary-coded form
        sval[idx] = k < len ? val[j0 + k] : static_cast<T>(0);
    }
}

// Sub-wavefront "vector" kernel: GL lanes per row, U rows per group in
// flight; lanes stride the row and the GL partial sums are tree-reduced with
// cross-lane shuffles.
template <typename T, bool ADD, bool DOT, int GL, int U>
__global__ __launch_bounds__(kBlock) void k_spmv_vector(
    int nrow, const int *__restrict__ row, const int *__restrict__ col, const T *__restrict__ val,
    const T *__restrict__ x, T *__restrict__ y, const T *__restrict__ w, double *part,
    const int *__restrict__ done) {
    if (done && *done) return;
    __shared__ double shd[kBlock / 64];
    constexpr int GPW = kBlock / GL;  // row groups per workgroup
    const int lane = threadIdx.x % GL, grp = threadIdx.x / GL;
    const Split sp = xcd_split(nrow, GPW * U);
    double dsum = 0;
    for (int base = sp.first; base < sp.hi; base += sp.step) {
        int rs[U], js[U], je[U];
        int maxlen = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            rs[u] = base + u * GPW + grp;
            const bool ok = rs[u] < sp.hi;
            js[u] = ok ? row[rs[u]] : 0;
            je[u] = ok ? row[rs[u] + 1] : 0;
            maxlen = max(maxlen, je[u] - js[u]);
        }
        T s[U];
#pragma unroll
        for (int u = 0; u < U; ++u) s[u] = 0;
        for (int k = lane; k < maxlen; k += GL) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int j = js[u] + k;
                if (j < je[u]) {
                    const int c = __builtin_nontemporal_load(&col[j]);
                    const T v = __builtin_nontemporal_load(&val[j]);
                    s[u] += v * x[c];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int off = GL / 2; off > 0; off >>= 1) s[u] += __shfl_xor(s[u], off, GL);
            if (lane == 0 && rs[u] < sp.hi) {
                T r = s[u];
                if (ADD) r = y[rs[u]] + r;
                y[rs[u]] = r;
                if (DOT) dsum += static_cast<double>(r) * static_cast<double>(w[rs[u]]);
            }
        }
    }
    if (DOT) {
        const double v[1] = {block_sum(dsum, shd)};
        store_partials<1>(v, part, 0, blockIdx.x);
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

// Row blocks: consecutive rows while the block holds <= tile nonzeros and
// <= max_rows rows; a longer row forms a block alone.
static std::vector<int> make_row_blocks(const std::vector<int> &row, int tile, int max_rows = kBlock) {
    const int n = static_cast<int>(row.size()) - 1;
    std::vector<int> rb;
    rb.reserve(n / 16 + 2);
    rb.push_back(0);
    int r = 0;
    while (r < n) {
        const int start = r;
        if (row[r + 1] - row[r] > tile) {
            r += 1;
        } else {
            while (r < n && r - start < max_rows && row[r + 1] - row[start] <= tile) ++r;
        }
        rb.push_back(r);
    }
    return rb;
}

// Runs f(lo, hi) over the parts of [0, n) on up to 16 host threads.
template <typename F>
static void host_parallel(int n, F f) {
    const unsigned hc = std::thread::hardware_concurrency();
    const int nt = n < 4096 ? 1 : static_cast<int>(std::min(16u, std::max(1u, hc)));
    if (nt == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back(f, static_cast<int>(static_cast<long long>(n) * t / nt),
                        static_cast<int>(static_cast<long long>(n) * (t + 1) / nt));
    for (auto &t : th) t.join();
}

// Largest LDS image of a slice in the LDS-staged layout (1024 elements: 8 KiB
// of FP64 x per wave), and its chunk size.
constexpr int kLdsMaxImage = 1024;
constexpr int kLdsChunk = 16;

struct LdsPlan {
    std::vector<int> cptr, cid;
    std::vector<unsigned short> code, dcode;
    int maxc = 0;
};

// Chunk lists (chunks of ch elements, ch a power of two) and 16-bit codes of
// the LDS-staged copy over the slot-pair layout with slice offsets ptr.  False
// when a slice's chunks exceed kLdsMaxImage elements.
static bool lds_plan(const hfh_csr *A, const std::vector<int> &ptr, int ch, LdsPlan &P) {
    int sh = 0;
    while ((1 << sh) < ch) ++sh;
    const int n = A->nrow, nsl = (n + 63) / 64;
    const int *row = A->h_row.data(), *col = A->h_col.data();
    std::vector<std::vector<int>> lists(nsl);
    host_parallel(nsl, [&](int lo, int hi) {
        std::vector<int> ids;
        for (int s = lo; s < hi; ++s) {
            ids.clear();
            for (int j = row[s * 64]; j < row[std::min(n, s * 64 + 64)]; ++j) ids.push_back(col[j] >> sh);
            std::sort(ids.begin(), ids.end());
            ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
            lists[s].assign(ids.begin(), ids.end());
        }
    });
    P.maxc = 0;
    P.cptr.assign(nsl + 1, 0);
    for (int s = 0; s < nsl; ++s) {
        P.maxc = std::max(P.maxc, static_cast<int>(lists[s].size()));
        P.cptr[s + 1] = P.cptr[s] + static_cast<int>(lists[s].size());
    }
    if (P.maxc * ch > kLdsMaxImage) return false;
    P.cid.resize(std::max(P.cptr[nsl], 1));
    P.code.assign(std::max(ptr[nsl], 1), 0);
    P.dcode.assign(std::max(n, 1), 0xffff);
    host_parallel(nsl, [&](int lo, int hi) {
        for (int s = lo; s < hi; ++s) {
            const std::vector<int> &ids = lists[s];
            std::copy(ids.begin(), ids.end(), P.cid.begin() + P.cptr[s]);
            const int base = ptr[s], L = (ptr[s + 1] - base) / 64, Pn = L & ~1;
            for (int lane = 0; lane < 64 && s * 64 + lane < n; ++lane) {
                const int r = s * 64 + lane, j0 = row[r], len = row[r + 1] - j0;
                for (int k = 0; k < len; ++k) {
                    const int c = col[j0 + k];
                    const int pos = static_cast<int>(std::lower_bound(ids.begin(), ids.end(), c >> sh) - ids.begin());
                    const unsigned short cd = static_cast<unsigned short>(pos * ch + (c & (ch - 1)));
                    const int idx = k < Pn ? base + (k >> 1) * 128 + 2 * lane + (k & 1) : base + k * 64 + lane;
                    P.code[idx] = cd;
                    if (c == r) P.dcode[r] = cd;
                }
            }
        }
    });
    return true;
}

// The sorted distinct column offsets col - row of A, offset 0 included.
// False when there are more than 255 (code 255 marks padding slots).  Rows
// whose offsets repeat the previous row's (most rows of a stencil matrix)
// are skipped; the row ranges run on host threads.
static bool dict_offsets(const hfh_csr *A, std::vector<int> &dict) {
    const int n = A->nrow;
    const int *row = A->h_row.data(), *col = A->h_col.data();
    constexpr int kMax = kPadCode;
    const int nt = n < 4096 ? 1 : static_cast<int>(std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
    std::vector<std::vector<long long>> found(nt);
    std::vector<char> over(nt, 0);
    host_parallel(nt, [&](int t0, int t1) {
        for (int t = t0; t < t1; ++t) {
            const int lo = static_cast<int>(static_cast<long long>(n) * t / nt);
            const int hi = static_cast<int>(static_cast<long long>(n) * (t + 1) / nt);
            std::vector<long long> &f = found[t];
            f.assign(1, 0);
            int prev = -1;
            for (int r = lo; r < hi && !over[t]; ++r) {
                const int j0 = row[r], len = row[r + 1] - j0;
                bool same = prev >= 0 && len == row[prev + 1] - row[prev];
                for (int k = 0; same && k < len; ++k)
                    same = col[j0 + k] - r == col[row[prev] + k] - prev;
                prev = r;
                if (same) continue;
                for (int k = 0; k < len; ++k) {
                    const long long o = static_cast<long long>(col[j0 + k]) - r;
                    auto it = std::lower_bound(f.begin(), f.end(), o);
                    if (it == f.end() || *it != o) {
                        f.insert(it, o);
                        if (static_cast<int>(f.size()) > kMax) {
                            over[t] = 1;
                            break;
                        }
                    }
                }
            }
        }
    });
    std::vector<long long> all;
    for (int t = 0; t < nt; ++t) {
        if (over[t]) return false;
        all.insert(all.end(), found[t].begin(), found[t].end());
    }
    std::sort(all.begin(), all.end());
    all.erase(std::unique(all.begin(), all.end()), all.end());
    if (static_cast<int>(all.size()) > kMax) return false;
    dict.assign(all.begin(), all.end());
    return true;
}

static void free_lds_copy(hfh_csr *A) {
    if (A->sx_code) (void)hipFree(A->sx_code);
    if (A->sx_dcode) (void)hipFree(A->sx_dcode);
    if (A->sx_cptr) (void)hipFree(A->sx_cptr);
    if (A->sx_cid) (void)hipFree(A->sx_cid);
    A->sx_code = A->sx_dcode = nullptr;
    A->sx_cptr = A->sx_cid = nullptr;
    A->sx_img = 0;
    A->sx_chunk = 0;
}

static int round8(long long g) { return static_cast<int>(g >= kNumXcd ? (g + kNumXcd - 1) / kNumXcd * kNumXcd : std::max(g, 1LL)); }

// Configures the SpMV kernel of A (host row pointer in A->h_row).
int spmv_configure(hfh_csr *A, int algo, int p1, int p2) {
    const int n = A->nrow;
    // buffers of the previous configuration may still be read by queued work
    HFH_HIP(hipStreamSynchronize(A->ctx->stream));
    if ((algo == HFH_SPMV_SELL || algo == HFH_SPMV_SELL_DICT || algo == HFH_SPMV_SELL_PAIRS ||
         algo == HFH_SPMV_SELL_LDS) && A->ncol == 0)
        algo = HFH_SPMV_STREAM;
    if (A->sd_dict) (void)hipFree(A->sd_dict);
    if (A->sd_code) (void)hipFree(A->sd_code);
    if (A->sd_wptr) (void)hipFree(A->sd_wptr);
    A->sd_dict = nullptr;
    A->sd_code = nullptr;
    A->sd_wptr = nullptr;
    A->sd_ndict = 0;
    free_lds_copy(A);
    const double mean = n ? static_cast<double>(A->nnz) / n : 0.0;
    // layout taken when the LDS-staged copy does not fit
    int fb_algo = HFH_SPMV_SELL_PAIRS, fb_p1 = 1;
    if (algo == 0) {  // automatic choice
        // SELL-64 unless padding to the longest row of each slice would add
        // more than a quarter to the stored entries; 4 column slots in flight
        // for long rows, 8 for short (stencil) rows.
        long long padded = 0;
        for (int sl = 0; sl < (n + 63) / 64; ++sl) {
            int L = 0;
            for (int r = sl * 64; r < std::min(n, sl * 64 + 64); ++r) L = std::max(L, A->h_row[r + 1] - A->h_row[r]);
            padded += 64LL * L;
        }
        if (A->ncol > 0 && static_cast<double>(padded) <= 1.25 * static_cast<double>(std::max(A->nnz, 1))) {
            // measured on gfx950: long rows run fastest with x staged in
            // LDS where the slices' x windows fit (3D Q1, 27 entries per row:
            // FP64 105 us against 125 us with slot pairs, FP32 64 us against
            // 83 us), else FP64 with slot pairs and FP32 with 4 slots in
            // flight.  Short (stencil) rows share few x entries within a
            // slice; they run fastest with one-byte dictionary codes where
            // the matrix has at most 255 distinct offsets col - row (7-point
            // FV 192^3: FP64 115 us against 129 us with plain SELL-64, FP32
            // 72 us against 79 us), else with SELL-64, 8 slots in flight
            if (mean >= 16.0) {
                algo = HFH_SPMV_SELL_LDS;
                p1 = 0;
                fb_algo = A->dtype == HFH_F64 ? HFH_SPMV_SELL_PAIRS : HFH_SPMV_SELL;
                fb_p1 = A->dtype == HFH_F64 ? 1 : 4;
            } else {
                algo = HFH_SPMV_SELL_DICT;
                p1 = A->dtype == HFH_F64 ? 1 : 2;   // FP32: 2 code words in flight (66 us against 71 us)
            }
        } else {
            algo = HFH_SPMV_STREAM;
            p1 = 2048;
        }
        p2 = 0;
    }
    // the dictionary of distinct column offsets (at most 255 codes)
    std::vector<int> dict;
    if (algo == HFH_SPMV_SELL_DICT && !dict_offsets(A, dict)) {
        algo = HFH_SPMV_SELL;
        p1 = 8;
    }
    const int cus = std::max(A->ctx->num_cu, 1);
    if (algo == HFH_SPMV_STREAM) {
        const int tile = p1 ? p1 : 2048;
        HFH_CHECK(tile == 512 || tile == 1024 || tile == 2048 || tile == 4096, HFH_ERR_INVALID,
                  "spmv: stream tile %d not in {512, 1024, 2048, 4096}", tile);
        std::vector<int> rb = make_row_blocks(A->h_row, tile);
        if (A->rblk) (void)hipFree(A->rblk);
        A->rblk = nullptr;
        int rc = dev_alloc(reinterpret_cast<void **>(&A->rblk), sizeof(int) * rb.size());
        if (rc) return rc;
        A->nrblk = static_cast<int>(rb.size()) - 1;
        HFH_HIP(hipMemcpy(A->rblk, rb.data(), sizeof(int) * rb.size(), hipMemcpyHostToDevice));
        const long long g = p2 > 0 ? std::min<long long>(p2, A->nrblk) : std::min<long long>(A->nrblk, static_cast<long long>(cus) * 16);
        A->spmv_algo = HFH_SPMV_STREAM;
        A->spmv_p1 = tile;
        A->nwg = round8(std::max<long long>(g, 1));
    } else if (algo == HFH_SPMV_SELL || algo == HFH_SPMV_SELL_PAIRS || algo == HFH_SPMV_SELL_LDS) {
        const int nsl = (n + 63) / 64;
        std::vector<int> ptr(nsl + 1, 0), len(std::max(n, 1), 0);
        long long off = 0;
        for (int sl = 0; sl < nsl; ++sl) {
            int L = 0;
            for (int r = sl * 64; r < std::min(n, sl * 64 + 64); ++r) {
                len[r] = A->h_row[r + 1] - A->h_row[r];
                L = std::max(L, len[r]);
            }
            ptr[sl] = static_cast<int>(off);
            off += 64LL * L;
            HFH_CHECK(off < (1LL << 31), HFH_ERR_UNSUPPORTED, "spmv: SELL copy exceeds 2^31 entries");
        }
        ptr[nsl] = static_cast<int>(off);
        LdsPlan plan;
        // 16-element chunks: 8-element chunks stage fewer unused entries but
        // measured no faster (3D Q1 FP64: 105.5 us against 104.7 us)
        const int ch = kLdsChunk;
        if (algo == HFH_SPMV_SELL_LDS && !lds_plan(A, ptr, ch, plan)) {
            algo = fb_algo;   // a slice's x window does not fit the LDS image
            p1 = fb_p1;
        }
        const bool lds = algo == HFH_SPMV_SELL_LDS, pairs = algo == HFH_SPMV_SELL_PAIRS;
        if (A->sl_ptr) (void)hipFree(A->sl_ptr);
        if (A->sl_len) (void)hipFree(A->sl_len);
        if (A->sl_col) (void)hipFree(A->sl_col);
        if (A->sl_val) (void)hipFree(A->sl_val);
        A->sl_ptr = A->sl_len = A->sl_col = nullptr;
        A->sl_val = nullptr;
        int rc;
        if ((rc = dev_alloc(reinterpret_cast<void **>(&A->sl_ptr), sizeof(int) * ptr.size())) ||
            (rc = dev_alloc(reinterpret_cast<void **>(&A->sl_len), sizeof(int) * len.size())) ||
            (!lds && (rc = dev_alloc(reinterpret_cast<void **>(&A->sl_col), sizeof(int) * static_cast<size_t>(std::max(off, 1LL))))) ||
            (rc = dev_alloc(&A->sl_val, A->elem() * static_cast<size_t>(std::max(off, 1LL)))))
            return rc;
        HFH_HIP(hipMemcpy(A->sl_ptr, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice));
        HFH_HIP(hipMemcpy(A->sl_len, len.data(), sizeof(int) * len.size(), hipMemcpyHostToDevice));
        if (lds) {
            if ((rc = dev_alloc(reinterpret_cast<void **>(&A->sx_code), sizeof(unsigned short) * plan.code.size())) ||
                (rc = dev_alloc(reinterpret_cast<void **>(&A->sx_dcode), sizeof(unsigned short) * plan.dcode.size())) ||
                (rc = dev_alloc(reinterpret_cast<void **>(&A->sx_cptr), sizeof(int) * plan.cptr.size())) ||
                (rc = dev_alloc(reinterpret_cast<void **>(&A->sx_cid), sizeof(int) * plan.cid.size())))
                return rc;
            HFH_HIP(hipMemcpy(A->sx_code, plan.code.data(), sizeof(unsigned short) * plan.code.size(), hipMemcpyHostToDevice));
            HFH_HIP(hipMemcpy(A->sx_dcode, plan.dcode.data(), sizeof(unsigned short) * plan.dcode.size(), hipMemcpyHostToDevice));
            HFH_HIP(hipMemcpy(A->sx_cptr, plan.cptr.data(), sizeof(int) * plan.cptr.size(), hipMemcpyHostToDevice));
            HFH_HIP(hipMemcpy(A->sx_cid, plan.cid.data(), sizeof(int) * plan.cid.size(), hipMemcpyHostToDevice));
            A->sx_img = plan.maxc * ch;
            A->sx_chunk = ch;
        }
        A->nslice = nsl;
        A->spmv_algo = algo;
        if (lds) A->spmv_p1 = 0;
        else if (pairs) A->spmv_p1 = p1 == 2 ? 2 : 1;
        else A->spmv_p1 = p1 == 2 || p1 == 8 || p1 == 16 ? p1 : 4;
        const long long need = (nsl + kBlock / 64 - 1) / (kBlock / 64);
        const long long g = p2 > 0 ? std::min<long long>(p2, need) : need;
        A->nwg = round8(std::max<long long>(g, 1));
        A->sell_dirty = true;  // filled from the device CSR by spmv_sell_fill
    } else if (algo == HFH_SPMV_SELL_DICT) {
        const int nsl = (n + 63) / 64;
        std::vector<int> ptr(nsl + 1, 0), wp(nsl + 1, 0), len(std::max(n, 1), 0);
        // values: each slice's width L of slots; code words: ceil(L / 4) per lane
        long long off = 0, woff = 0;
        for (int sl = 0; sl < nsl; ++sl) {
            int L = 0;
            for (int r = sl * 64; r < std::min(n, sl * 64 + 64); ++r) {
                len[r] = A->h_row[r + 1] - A->h_row[r];
                L = std::max(L, len[r]);
            }
            ptr[sl] = static_cast<int>(off);
            wp[sl] = static_cast<int>(woff);
            off += 64LL * L;
            woff += 64LL * ((L + 3) / 4);
            HFH_CHECK(off < (1LL << 31), HFH_ERR_UNSUPPORTED, "spmv: SELL copy exceeds 2^31 entries");
        }
        ptr[nsl] = static_cast<int>(off);
        wp[nsl] = static_cast<int>(woff);
        // code words: byte t of word w of a lane = code of slot 4w + t
        std::vector<unsigned> words(std::max(woff, 1LL), 0u);
        // padding slots take kPadCode (offset 0 in the kernel's table: their
        // gather reads x[min(r, ncol - 1)], in range); zero: the diagonal's
        const unsigned zero = static_cast<unsigned>(std::lower_bound(dict.begin(), dict.end(), 0) - dict.begin());
        host_parallel(nsl, [&](int s0, int s1) {
            std::vector<unsigned> codes;
            int prev = -1;   // previous row whose codes are in `codes`
            for (int sl = s0; sl < s1; ++sl) {
                const int L4 = (wp[sl + 1] - wp[sl]) / 64 * 4;   // the slots the code words cover
                for (int lane = 0; lane < 64; ++lane) {
                    const int r = sl * 64 + lane;
                    const int l = r < n ? len[r] : 0, j0 = r < n ? A->h_row[r] : 0;
                    // a row with the previous row's offsets has its codes
                    bool same = prev >= 0 && r < n && l == static_cast<int>(codes.size());
                    for (int k = 0; same && k < l; ++k)
                        same = A->h_col[j0 + k] - r == A->h_col[A->h_row[prev] + k] - prev;
                    if (!same && r < n) {
                        codes.resize(l);
                        for (int k = 0; k < l; ++k)
                            codes[k] = static_cast<unsigned>(
                                std::lower_bound(dict.begin(), dict.end(), A->h_col[j0 + k] - r) - dict.begin());
                    }
                    if (r < n) prev = r;
                    for (int k = 0; k < L4; ++k) {
                        const unsigned code = k < l ? codes[k] : kPadCode;
                        words[wp[sl] + (k / 4) * 64 + lane] |= code << (8 * (k % 4));
                    }
                }
            }
        });
        if (A->sl_ptr) (void)hipFree(A->sl_ptr);
        if (A->sl_len) (void)hipFree(A->sl_len);
        if (A->sl_col) (void)hipFree(A->sl_col);
        if (A->sl_val) (void)hipFree(A->sl_val);
        A->sl_ptr = A->sl_len = A->sl_col = nullptr;
        A->sl_val = nullptr;
        int rc;
        if ((rc = dev_alloc(reinterpret_cast<void **>(&A->sl_ptr), sizeof(int) * ptr.size())) ||
            (rc = dev_alloc(reinterpret_cast<void **>(&A->sl_len), sizeof(int) * len.size())) ||
            (rc = dev_alloc(&A->sl_val, A->elem() * static_cast<size_t>(std::max(off, 1LL)))) ||
            (rc = dev_alloc(reinterpret_cast<void **>(&A->sd_wptr), sizeof(int) * wp.size())) ||
            (rc = dev_alloc(reinterpret_cast<void **>(&A->sd_code), sizeof(unsigned) * words.size())) ||
            (rc = dev_alloc(reinterpret_cast<void **>(&A->sd_dict), sizeof(int) * dict.size())))
            return rc;
        HFH_HIP(hipMemcpy(A->sl_ptr, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice));
        HFH_HIP(hipMemcpy(A->sl_len, len.data(), sizeof(int) * len.size(), hipMemcpyHostToDevice));
        HFH_HIP(hipMemcpy(A->sd_wptr, wp.data(), sizeof(int) * wp.size(), hipMemcpyHostToDevice));
        HFH_HIP(hipMemcpy(A->sd_code, words.data(), sizeof(unsigned) * words.size(), hipMemcpyHostToDevice));
        HFH_HIP(hipMemcpy(A->sd_dict, dict.data(), sizeof(int) * dict.size(), hipMemcpyHostToDevice));
        A->sd_ndict = static_cast<int>(dict.size());
        A->sd_code0 = static_cast<int>(zero);
        A->nslice = nsl;
        A->spmv_algo = HFH_SPMV_SELL_DICT;
        A->spmv_p1 = p1 == 2 ? 2 : 1;
        const long long need = (nsl + kBlock / 64 - 1) / (kBlock / 64);
        // the kernel's look-ahead follows the same choice (kDictCappedGrid)
        const bool capped = A->dtype == HFH_F32 ? kDictCappedGrid<float> : kDictCappedGrid<double>;
        const long long dflt = capped ? std::min<long long>(need, static_cast<long long>(cus) * 8) : need;
        const long long g = p2 > 0 ? std::min<long long>(p2, need) : dflt;
        A->nwg = round8(std::max<long long>(g, 1));
        A->sell_dirty = true;  // values filled from the device CSR by spmv_sell_fill
    } else if (algo == HFH_SPMV_WAVE) {
        const int tile = p1 ? p1 : 512;
        HFH_CHECK(tile == 256 || tile == 512 || tile == 1024, HFH_ERR_INVALID,
                  "spmv: wave tile %d not in {256, 512, 1024}", tile);
        std::vector<int> rb = make_row_blocks(A->h_row, tile, 63);  // 63 rows: 64 row pointers, one per lane
        if (A->rblk) (void)hipFree(A->rblk);
        A->rblk = nullptr;
        int rc = dev_alloc(reinterpret_cast<void **>(&A->rblk), sizeof(int) * rb.size());
        if (rc) return rc;
        A->nrblk = static_cast<int>(rb.size()) - 1;
        HFH_HIP(hipMemcpy(A->rblk, rb.data(), sizeof(int) * rb.size(), hipMemcpyHostToDevice));
        const long long need = (A->nrblk + kBlock / 64 - 1) / (kBlock / 64);
        const long long g = p2 > 0 ? std::min<long long>(p2, need) : std::min<long long>(need, static_cast<long long>(cus) * 8);
        A->spmv_algo = HFH_SPMV_WAVE;
        A->spmv_p1 = tile;
        A->nwg = round8(std::max<long long>(g, 1));
    } else if (algo == HFH_SPMV_VECTOR) {
        int gl = p1;
        if (gl == 0) {
            gl = 4;
            while (gl < 64 && gl < mean) gl *= 2;
        }
        HFH_CHECK(gl == 4 || gl == 8 || gl == 16 || gl == 32 || gl == 64, HFH_ERR_INVALID,
                  "spmv: vector group %d not in {4, 8, 16, 32, 64}", gl);
        const long long rows_per_wg = (kBlock / gl) * 2;
        const long long need = (n + rows_per_wg - 1) / rows_per_wg;
        const long long g = p2 > 0 ? p2 : std::min<long long>(need, static_cast<long long>(cus) * 16);
        A->spmv_algo = HFH_SPMV_VECTOR;
        A->spmv_p1 = gl;
        A->nwg = round8(std::max<long long>(g, 1));
    } else {
        set_error("spmv: unknown algorithm %d", algo);
        return HFH_ERR_INVALID;
    }
    return HFH_OK;
}

// The LDS-staged kernel with the image sized to the matrix's largest chunk list.
template <typename T, bool ADD, bool DOT, int JAC, int CH>
static void lds_launch_img(const hfh_csr *A, dim3 g, hipStream_t s, int xal, const T *x, T *y, const T *w, double *part,
                           const int *done, const T *b, const T *invd, T jw, T jomw) {
    const dim3 bl(kBlock);
    const T *sv = static_cast<const T *>(A->sl_val);
    const size_t lds = sizeof(T) * (kBlock / 64) * static_cast<size_t>(A->sx_img);
    if (A->sx_img <= 256)
        hipLaunchKernelGGL((k_spmv_sellx<T, ADD, DOT, JAC, CH, 256>), g, bl, lds, s, A->nrow, A->ncol, A->nslice, A->sl_ptr, A->sl_len, A->sx_code, sv, A->sx_cptr, A->sx_cid, A->sx_img, xal, x, y, w, part, done, A->sx_dcode, b, invd, jw, jomw);
    else if (A->sx_img <= 512)
        hipLaunchKernelGGL((k_spmv_sellx<T, ADD, DOT, JAC, CH, 512>), g, bl, lds, s, A->nrow, A->ncol, A->nslice, A->sl_ptr, A->sl_len, A->sx_code, sv, A->sx_cptr, A->sx_cid, A->sx_img, xal, x, y, w, part, done, A->sx_dcode, b, invd, jw, jomw);
    else if (A->sx_img <= 768)
        hipLaunchKernelGGL((k_spmv_sellx<T, ADD, DOT, JAC, CH, 768>), g, bl, lds, s, A->nrow, A->ncol, A->nslice, A->sl_ptr, A->sl_len, A->sx_code, sv, A->sx_cptr, A->sx_cid, A->sx_img, xal, x, y, w, part, done, A->sx_dcode, b, invd, jw, jomw);
    else
        hipLaunchKernelGGL((k_spmv_sellx<T, ADD, DOT, JAC, CH, kLdsMaxImage>), g, bl, lds, s, A->nrow, A->ncol, A->nslice, A->sl_ptr, A->sl_len, A->sx_code, sv, A->sx_cptr, A->sx_cid, A->sx_img, xal, x, y, w, part, done, A->sx_dcode, b, invd, jw, jomw);
}

template <typename T, bool ADD, bool DOT, int JAC>
static void lds_launch(const hfh_csr *A, int nwg, const T *x, T *y, const T *w, double *part, const int *done,
                       const T *b, const T *invd, T jw, T jomw) {
    hipStream_t s = A->ctx->stream;
    const dim3 g(nwg);
    const int xal = (reinterpret_cast<uintptr_t>(x) & 15) == 0;
    lds_launch_img<T, ADD, DOT, JAC, kLdsChunk>(A, g, s, xal, x, y, w, part, done, b, invd, jw, jomw);
}

// The dictionary-coded kernel; code0 is the code of offset 0 (the diagonal).
template <typename T, bool ADD, bool DOT, int JAC>
static void dict_launch(const hfh_csr *A, int nwg, const T *x, T *y, const T *w, double *part, const int *done,
                        const T *b, const T *invd, T jw, T jomw) {
    hipStream_t s = A->ctx->stream;
    const dim3 g(nwg), bl(kBlock);
    const T *sv = static_cast<const T *>(A->sl_val);
    const unsigned code0 = static_cast<unsigned>(A->sd_code0);
    if (A->spmv_p1 == 2)
        hipLaunchKernelGGL((k_spmv_selld<T, ADD, DOT, 2, JAC>), g, bl, 0, s, A->nrow, A->ncol, A->nslice, A->sl_ptr, A->sd_wptr, A->sd_code, A->sd_dict, A->sd_ndict, sv, x, y, w, part, done, code0, b, invd, jw, jomw);
    else
        hipLaunchKernelGGL((k_spmv_selld<T, ADD, DOT, 1, JAC>), g, bl, 0, s, A->nrow, A->ncol, A->nslice, A->sl_ptr, A->sd_wptr, A->sd_code, A->sd_dict, A->sd_ndict, sv, x, y, w, part, done, code0, b, invd, jw, jomw);
}

template <typename T>
int spmv_coded_jacobi(const hfh_csr *A, const T *x, T *xn, const T *b, const T *invd, bool damped, T w, T omw) {
    const bool lds = A->spmv_algo == HFH_SPMV_SELL_LDS;
    HFH_CHECK(lds || A->spmv_algo == HFH_SPMV_SELL_DICT, HFH_ERR_INVALID,
              "spmv_coded_jacobi: matrix has no LDS-staged or dictionary-coded copy");
    if (A->nrow == 0) return HFH_OK;
    const int need = std::max((A->nslice + kBlock / 64 - 1) / (kBlock / 64), 1);
    if (lds && damped)
        lds_launch<T, false, false, 2>(A, need, x, xn, nullptr, nullptr, nullptr, b, invd, w, omw);
    else if (lds)
        lds_launch<T, false, false, 1>(A, need, x, xn, nullptr, nullptr, nullptr, b, invd, w, omw);
    else if (damped)
        dict_launch<T, false, false, 2>(A, need, x, xn, nullptr, nullptr, nullptr, b, invd, w, omw);
    else
        dict_launch<T, false, false, 1>(A, need, x, xn, nullptr, nullptr, nullptr, b, invd, w, omw);
    return launch_check(lds ? "k_spmv_sellx (Jacobi)" : "k_spmv_selld (Jacobi)");
}
template int spmv_coded_jacobi<double>(const hfh_csr *, const double *, double *, const double *, const double *,
                                       bool, double, double);
template int spmv_coded_jacobi<float>(const hfh_csr *, const float *, float *, const float *, const float *, bool,
                                      float, float);

template <typename T, bool ADD, bool DOT>
static void launch_t(const hfh_csr *A, const void *x, void *y, const void *w, double *part, const int *done) {
    hipStream_t s = A->ctx->stream;
    const dim3 g(A->nwg), bl(kBlock);
    const T *xv = (const T *)x, *wv = (const T *)w, *val = (const T *)A->val;
    T *yv = (T *)y;
    if (A->spmv_algo == HFH_SPMV_SELL_LDS) {
        lds_launch<T, ADD, DOT, 0>(A, A->nwg, xv, yv, wv, part, done, nullptr, nullptr, T(0), T(0));
    } else if (A->spmv_algo == HFH_SPMV_SELL_PAIRS) {
        if (A->spmv_p1 == 2)
            hipLaunchKernelGGL((k_spmv_sellp<T, ADD, DOT, 2>), g, bl, 0, s, A->nrow, A->nslice, A->sl_ptr, A->sl_len, A->sl_col, (const T *)A->sl_val, xv, yv, wv, part, done);
        else
            hipLaunchKernelGGL((k_spmv_sellp<T, ADD, DOT, 1>), g, bl, 0, s, A->nrow, A->nslice, A->sl_ptr, A->sl_len, A->sl_col, (const T *)A->sl_val, xv, yv, wv, part, done);
    } else if (A->spmv_algo == HFH_SPMV_SELL_DICT) {
        dict_launch<T, ADD, DOT, 0>(A, A->nwg, xv, yv, wv, part, done, nullptr, nullptr, T(0), T(0));
    } else if (A->spmv_algo == HFH_SPMV_SELL) {
        switch (A->spmv_p1) {
        case 2: hipLaunchKernelGGL((k_spmv_sell<T, ADD, DOT, 2>), g, bl, 0, s, A->nrow, A->nslice, A->sl_ptr, A->sl_len, A->sl_col, (const T *)A->sl_val, xv, yv, wv, part, done); break;
        case 8: hipLaunchKernelGGL((k_spmv_sell<T, ADD, DOT, 8>), g, bl, 0, s, A->nrow, A->nslice, A->sl_ptr, A->sl_len, A->sl_col, (const T *)A->sl_val, xv, yv, wv, part, done); break;
        case 16: hipLaunchKernelGGL((k_spmv_sell<T, ADD, DOT, 16>), g, bl, 0, s, A->nrow, A->nslice, A->sl_ptr, A->sl_len, A->sl_col, (const T *)A->sl_val, xv, yv, wv, part, done); break;
        default: hipLaunchKernelGGL((k_spmv_sell<T, ADD, DOT, 4>), g, bl, 0, s, A->nrow, A->nslice, A->sl_ptr, A->sl_len, A->sl_col, (const T *)A->sl_val, xv, yv, wv, part, done); break;
        }
    } else if (A->spmv_algo == HFH_SPMV_WAVE) {
        switch (A->spmv_p1) {
        case 256: hipLaunchKernelGGL((k_spmv_wave<T, ADD, DOT, 256>), g, bl, 0, s, A->row, A->col, val, xv, yv, A->rblk, A->nrblk, wv, part, done); break;
        case 512: hipLaunchKernelGGL((k_spmv_wave<T, ADD, DOT, 512>), g, bl, 0, s, A->row, A->col, val, xv, yv, A->rblk, A->nrblk, wv, part, done); break;
        default: hipLaunchKernelGGL((k_spmv_wave<T, ADD, DOT, 1024>), g, bl, 0, s, A->row, A->col, val, xv, yv, A->rblk, A->nrblk, wv, part, done); break;
        }
    } else if (A->spmv_algo == HFH_SPMV_STREAM) {
        switch (A->spmv_p1) {
        case 512: hipLaunchKernelGGL((k_spmv_stream<T, ADD, DOT, 512>), g, bl, 0, s, A->row, A->col, val, xv, yv, A->rblk, A->nrblk, wv, part, done); break;
        case 1024: hipLaunchKernelGGL((k_spmv_stream<T, ADD, DOT, 1024>), g, bl, 0, s, A->row, A->col, val, xv, yv, A->rblk, A->nrblk, wv, part, done); break;
        case 2048: hipLaunchKernelGGL((k_spmv_stream<T, ADD, DOT, 2048>), g, bl, 0, s, A->row, A->col, val, xv, yv, A->rblk, A->nrblk, wv, part, done); break;
        default: hipLaunchKernelGGL((k_spmv_stream<T, ADD, DOT, 4096>), g, bl, 0, s, A->row, A->col, val, xv, yv, A->rblk, A->nrblk, wv, part, done); break;
        }
    } else {
        switch (A->spmv_p1) {
        case 4: hipLaunchKernelGGL((k_spmv_vector<T, ADD, DOT, 4, 2>), g, bl, 0, s, A->nrow, A->row, A->col, val, xv, yv, wv, part, done); break;
        case 8: hipLaunchKernelGGL((k_spmv_vector<T, ADD, DOT, 8, 2>), g, bl, 0, s, A->nrow, A->row, A->col, val, xv, yv, wv, part, done); break;
        case 16: hipLaunchKernelGGL((k_spmv_vector<T, ADD, DOT, 16, 2>), g, bl, 0, s, A->nrow, A->row, A->col, val, xv, yv, wv, part, done); break;
        case 32: hipLaunchKernelGGL((k_spmv_vector<T, ADD, DOT, 32, 2>), g, bl, 0, s, A->nrow, A->row, A->col, val, xv, yv, wv, part, done); break;
        default: hipLaunchKernelGGL((k_spmv_vector<T, ADD, DOT, 64, 2>), g, bl, 0, s, A->nrow, A->row, A->col, val, xv, yv, wv, part, done); break;
        }
    }
}

// Builds the SELL-64 values and columns from the device CSR (after an
// upload, a cast or a change of algorithm).
int spmv_sell_fill(hfh_csr *A) {
    if ((A->spmv_algo != HFH_SPMV_SELL && A->spmv_algo != HFH_SPMV_SELL_DICT && A->spmv_algo != HFH_SPMV_SELL_PAIRS &&
         A->spmv_algo != HFH_SPMV_SELL_LDS) ||
        !A->sell_dirty || A->nrow == 0)
        return HFH_OK;
    const int pairs = A->spmv_algo == HFH_SPMV_SELL_PAIRS || A->spmv_algo == HFH_SPMV_SELL_LDS;
    const dim3 g((A->nslice * 64 + kBlock - 1) / kBlock);
    if (A->dtype == HFH_F64)
        hipLaunchKernelGGL(k_csr_to_sell<double>, g, dim3(kBlock), 0, A->ctx->stream, A->nrow, A->ncol, A->nslice,
                           A->row, A->col, (const double *)A->val, A->sl_ptr, A->sl_col, (double *)A->sl_val,
                           pairs);
    else
        hipLaunchKernelGGL(k_csr_to_sell<float>, g, dim3(kBlock), 0, A->ctx->stream, A->nrow, A->ncol, A->nslice,
                           A->row, A->col, (const float *)A->val, A->sl_ptr, A->sl_col, (float *)A->sl_val, pairs);
    int rc = launch_check("k_csr_to_sell");
    if (rc) return rc;
    A->sell_dirty = false;
    return HFH_OK;
}

int spmv_launch_ex(const hfh_csr *A, const void *x, void *y, int mode, const void *w, double *part,
                   const int *done) {
    if (A->nrow == 0) return HFH_OK;
    const bool dot = w != nullptr;
    if (A->dtype == HFH_F64) {
        if (mode == 0 && !dot) launch_t<double, false, false>(A, x, y, w, part, done);
        else if (mode == 0) launch_t<double, false, true>(A, x, y, w, part, done);
        else if (!dot) launch_t<double, true, false>(A, x, y, w, part, done);
        else launch_t<double, true, true>(A, x, y, w, part, done);
    } else {
        if (mode == 0 && !dot) launch_t<float, false, false>(A, x, y, w, part, done);
        else if (mode == 0) launch_t<float, false, true>(A, x, y, w, part, done);
        else if (!dot) launch_t<float, true, false>(A, x, y, w, part, done);
        else launch_t<float, true, true>(A, x, y, w, part, done);
    }
    return launch_check("k_spmv");
}

int spmv_launch(const hfh_csr *A, const void *x, void *y, int mode) {
    ScopedKernelTimer t(A->ctx, 0);
    return spmv_launch_ex(A, x, y, mode, nullptr, nullptr, nullptr);
}

}  // namespace hfh

using namespace hfh;

extern "C" int hfh_csr_create(hfh_ctx *ctx, int nrow, int ncol, int nnz, int dtype, hfh_csr **out) {
    HFH_CHECK(ctx && out, HFH_ERR_INVALID, "hfh_csr_create: null pointer");
    HFH_CHECK(nrow >= 0 && ncol >= 0 && nnz >= 0, HFH_ERR_INVALID, "hfh_csr_create: negative size");
    HFH_CHECK(dtype == HFH_F64 || dtype == HFH_F32, HFH_ERR_INVALID, "hfh_csr_create: bad dtype %d", dtype);
    hfh_csr *A = new hfh_csr();
    A->ctx = ctx;
    ++ctx->live;
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
    if (A->ctx) --A->ctx->live;
    if (A->row) (void)hipFree(A->row);
    if (A->col) (void)hipFree(A->col);
    if (A->val) (void)hipFree(A->val);
    if (A->rblk) (void)hipFree(A->rblk);
    if (A->diag) (void)hipFree(A->diag);
    if (A->sl_ptr) (void)hipFree(A->sl_ptr);
    if (A->sl_len) (void)hipFree(A->sl_len);
    if (A->sl_col) (void)hipFree(A->sl_col);
    if (A->sl_val) (void)hipFree(A->sl_val);
    if (A->sd_dict) (void)hipFree(A->sd_dict);
    if (A->sd_code) (void)hipFree(A->sd_code);
    if (A->sd_wptr) (void)hipFree(A->sd_wptr);
    free_lds_copy(A);
    delete A;
    return HFH_OK;
}

extern "C" int hfh_csr_get_spmv(const hfh_csr *A, int *algo, int *param1) {
    HFH_CHECK(A, HFH_ERR_INVALID, "hfh_csr_get_spmv: null matrix");
    if (algo) *algo = A->spmv_algo;
    if (param1) *param1 = A->spmv_p1;
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
    if (A->nnz) A->h_col.assign(col, col + A->nnz);
    int rc = spmv_configure(A, A->spmv_algo_req, A->spmv_p1_req, A->spmv_p2_req);
    if (rc) return rc;
    HFH_HIP(hipMemcpyAsync(A->row, row, sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    if (A->nnz) {
        HFH_HIP(hipMemcpyAsync(A->col, col, sizeof(int) * static_cast<size_t>(A->nnz), hipMemcpyHostToDevice, s));
        HFH_HIP(hipMemcpyAsync(A->val, val, A->elem() * static_cast<size_t>(A->nnz), hipMemcpyHostToDevice, s));
    }
    if (n > 0) {
        hipLaunchKernelGGL(k_find_diag<int>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, A->row, A->col, A->diag);
        if ((rc = launch_check("k_find_diag"))) return rc;
    }
    A->has_structure = true;
    A->sorted = true;
    ++A->version;
    A->sell_dirty = true;
    if ((rc = spmv_sell_fill(A))) return rc;
    HFH_HIP(hipStreamSynchronize(s));
    return HFH_OK;
}

int hfh::spmv_refresh_structure(hfh_csr *A) {
    hipStream_t s = A->ctx->stream;
    int rc = spmv_configure(A, A->spmv_algo_req, A->spmv_p1_req, A->spmv_p2_req);
    if (rc) return rc;
    if (A->nrow > 0) {
        hipLaunchKernelGGL(k_find_diag<int>, dim3((A->nrow + kBlock - 1) / kBlock), dim3(kBlock), 0, s, A->nrow,
                           A->row, A->col, A->diag);
        if ((rc = launch_check("k_find_diag"))) return rc;
    }
    ++A->version;
    A->sell_dirty = true;
    return spmv_sell_fill(A);
}

extern "C" int hfh_csr_upload_values(hfh_csr *A, const void *val) {
    HFH_CHECK(A && (val || A->nnz == 0), HFH_ERR_INVALID, "hfh_csr_upload_values: null pointer");
    HFH_CHECK(A->has_structure, HFH_ERR_INVALID, "hfh_csr_upload_values: structure not set");
    if (A->nnz == 0) return HFH_OK;
    HFH_HIP(hipMemcpyAsync(A->val, val, A->elem() * static_cast<size_t>(A->nnz), hipMemcpyHostToDevice, A->ctx->stream));
    ++A->version;
    A->sell_dirty = true;
    int rc = spmv_sell_fill(A);
    if (rc) return rc;
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
    dst->h_col = src->h_col;
    dst->sorted = src->sorted;
    // the destination's requested layout; else the source's requested one;
    // else the automatic choice for the destination's value type
    const bool own = dst->spmv_algo_req != 0;
    int rc = spmv_configure(dst, own ? dst->spmv_algo_req : src->spmv_algo_req,
                            own ? dst->spmv_p1_req : src->spmv_p1_req, own ? dst->spmv_p2_req : src->spmv_p2_req);
    if (rc) return rc;
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
    dst->has_structure = true;
    ++dst->version;
    dst->sell_dirty = true;
    if ((rc = spmv_sell_fill(dst))) return rc;
    HFH_HIP(hipStreamSynchronize(s));
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

extern "C" int hfh_csr_set_spmv(hfh_csr *A, int algo, int p1, int p2) {
    HFH_CHECK(A, HFH_ERR_INVALID, "hfh_csr_set_spmv: null matrix");
    HFH_CHECK(algo >= 0 && algo <= HFH_SPMV_SELL_LDS && p2 >= 0, HFH_ERR_INVALID, "hfh_csr_set_spmv: bad argument");
    A->spmv_algo_req = algo;
    A->spmv_p1_req = p1;
    A->spmv_p2_req = p2;
    if (!A->has_structure) return HFH_OK;
    HFH_HIP(hipStreamSynchronize(A->ctx->stream));
    int rc = spmv_configure(A, algo, p1, p2);
    if (rc) return rc;
    if ((rc = spmv_sell_fill(A))) return rc;
    HFH_HIP(hipStreamSynchronize(A->ctx->stream));
    return HFH_OK;
}
