// hfh_batch.hip -- serving engine: many independent SPD systems solved
// together by FP32-inner / FP64-outer CG with continuous admission.
//
// Each admitted system is solved by defect correction in the form of
// MPIR<D,S>::Solve (linear_solver/mpir.cc:69-327): FP64 true residual
// r = b - A x, FP32 inner Jacobi-preconditioned CG on A c = r
// (preconditioner lpreconditioner.cc:104-166), FP64 update x += c, until the
// true residual is at most max(atol, rtol * initial residual).  The engine
// has no reference counterpart; it runs the same algorithm for many systems
// at once.  The inner CG is the Chronopoulos-Gear form of CG::SolvePrecond
// (linear_solver/cg.cc:84-248): the same iterates in exact arithmetic, with
// both inner products of a step formed after its one SpMV, so a step needs
// one grid-wide reduction instead of two:
//   p = u + beta p,  s = w + beta s,  c += alpha p,  r -= alpha s,
//   u = M r,  w = A u,  gamma' = r.u,  delta = w.u,
//   beta = gamma' / gamma,  alpha = gamma' / (delta - beta gamma' / alpha)
// (w = A u and s = A p are carried instead of the second product A p).
//
// Device layout.  One arena of `hbm_budget` bytes holds every admitted
// system in a contiguous region (first-fit, 256-byte aligned sections):
//   sl_ptr[nslice+1] sl_len[n] cols v64[E] v32[E] invd[n] b[n] x[n]   (uploaded / assembled)
//   r64[n] r[n] w[n] s[n] u[n] c[n] p[n] part[3*ntiles]                  (work space)
// with the matrix in SELL-64 form (slices of 64 rows, column-major, E
// padded entries; FP64 values for the true residual, FP32 values for the
// inner solve, one shared column array).  `cols` is either sl_col[E] or,
// for every FV system (5 or 7 distinct offsets col - row), one-byte codes
// into the system's offset table dict[256]: code word w of a lane holds
// slots 4w..4w+3, ceil(L/4) words per lane of a slice of width L, so the
// values keep the slice width; a row's padding slots hold kPadCode.  A slot table holds
// each system's descriptor (pointers, sizes) and its solver state.  A tile
// table lists every active system's tiles (16 slices = 1024 rows each).
//
// One step of all systems is three launches:
//   K1 tiles : inner  p = u + beta p, s = w + beta s, c += alpha p,
//                     r -= alpha s, u = invd r (stored for K2's gathers),
//                     partials r.u and r.r
//              setup  r = fl(r64), u = invd r and the partials
//              outer  x += c, c = 0 (after an inner solve; c = 0 first)
//   K2 tiles : inner / setup  w = A32 u, partial w.u
//              outer  r64 = b - A64 x, partial r64.r64
//   K3 slots : the decision from the system's partials: alpha and beta, the
//              inner stop test, the outer stop test, the next phase
// The tile kernels loop over the tile table with a fixed grid, and the slot
// kernel over the slot table, so `chunk` steps are captured once in a
// hipGraph and replayed while systems come and go: admission and retirement
// only rewrite the tables between replays.  A system's arithmetic depends
// only on its own data and its own fixed tiling (partials summed per system
// in tile order), so its result is bit-identical alone and in any mix.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "hfh_internal.h"

namespace hfh {
namespace {

constexpr int kTileSlices = 16;     // 1024 rows per tile
constexpr int kTileRows = kTileSlices * 64;
// Workgroups of the tile kernels: eight times the resident ones, so the
// hardware hands out tiles as workgroups finish (a fixed stride over many
// tiles leaves the slowest workgroups last: closed-loop capacity 142 -> 159
// solves/s from 2,048 to 8,192; the workgroups past the last tile end at
// once).  Tiles of 1,024 rows give a lightly loaded step twice the
// workgroups of 2,048-row tiles and half the slices per wave (0.59 -> 0.49
// ms per 16-step replay at 116 arrivals/s), at the same capacity; 512-row
// tiles were slower again (0.52 ms, 0.66 ms on 32,768 workgroups).
constexpr int kTileGrid = 16384;
constexpr int kSellU = 8;           // column slots per SELL step
constexpr size_t kAlign = 256;
constexpr int kParts = 3;           // partials per tile: r.u, w.u, r.r (outer: 0, 0, r64.r64)

// solver phases (SysState::phase): what the next K1 and K2 do
enum : int {
    P_IDLE = 0,     // free slot
    P_FIRST = 1,    // c = 0; the initial true residual
    P_CORRECT = 2,  // x += c, c = 0; the true residual
    P_SETUP = 3,    // the start of an inner solve from r64: r, u, w = A u
    P_START = 4,    // first inner step (p = u, s = w)
    P_GO = 5,       // inner step with beta
    P_DONE = 7      // finished (state holds the outcome)
};

struct SysDesc {
    int n, nslice, ntiles, ndict;   // ndict > 0: coded columns
    int small;                      // solved by kb_small, not by the tile and slot kernels
    const int *sl_ptr;
    const int *sl_len;
    const int *sl_col;
    const unsigned *sl_code;
    const int *sl_wptr;
    const int *dict;
    const double *v64;
    const float *v32;
    const float *invd;
    const double *b;
    double *x;
    double *r64;
    float *r, *w, *s, *u, *c, *p;
    double *part;
};

// Device-side view of a descriptor with its pointers in the global address
// space: pointers loaded from a descriptor in memory are otherwise generic,
// and generic (flat) loads make every wait drain the LDS counter as well.
template <typename T>
using gptr = __attribute__((address_space(1))) T *;

struct GDesc {
    int n, nslice, ntiles, ndict;
    gptr<const int> sl_ptr, sl_len, sl_col;
    gptr<const unsigned> sl_code;
    gptr<const int> sl_wptr, dict;
    gptr<const double> v64;
    gptr<const float> v32, invd;
    gptr<const double> b;
    gptr<double> x, r64;
    gptr<float> r, w, s, u, c, p;
    gptr<double> part;
};

__device__ __forceinline__ GDesc gview(const SysDesc &d) {
    GDesc g;
    g.n = d.n;
    g.nslice = d.nslice;
    g.ntiles = d.ntiles;
    g.ndict = d.ndict;
    g.sl_ptr = (gptr<const int>)d.sl_ptr;
    g.sl_len = (gptr<const int>)d.sl_len;
    g.sl_col = (gptr<const int>)d.sl_col;
    g.sl_code = (gptr<const unsigned>)d.sl_code;
    g.sl_wptr = (gptr<const int>)d.sl_wptr;
    g.dict = (gptr<const int>)d.dict;
    g.v64 = (gptr<const double>)d.v64;
    g.v32 = (gptr<const float>)d.v32;
    g.invd = (gptr<const float>)d.invd;
    g.b = (gptr<const double>)d.b;
    g.x = (gptr<double>)d.x;
    g.r64 = (gptr<double>)d.r64;
    g.r = (gptr<float>)d.r;
    g.w = (gptr<float>)d.w;
    g.s = (gptr<float>)d.s;
    g.u = (gptr<float>)d.u;
    g.c = (gptr<float>)d.c;
    g.p = (gptr<float>)d.p;
    g.part = (gptr<double>)d.part;
    return g;
}

struct SysState {
    int phase, status, outer, inner, inner_total;
    int id;   // request id of the occupant (the kernels never change it)
    double res0, res, tol, rr0, alpha, beta, gamma;
};

struct Tile {
    int sys, s0, ns, lt;   // slot, first slice, slices, tile index within the system
};

struct KParams {
    double rtol, atol, inner_rtol;
    int outer_max, inner_max;
};

// Products and sums of the SELL row sums are rounded separately (no
// contraction), in ascending column order, as the CPU row loop does.
template <typename T>
__device__ __forceinline__ T bmul(T a, T b) {
#pragma clang fp contract(off)
    return a * b;
}
template <typename T>
__device__ __forceinline__ T badd(T a, T b) {
#pragma clang fp contract(off)
    return a + b;
}

typedef float bf4 __attribute__((ext_vector_type(4)));
typedef double bd2 __attribute__((ext_vector_type(2)));

// The SpMV operand at row j.
template <typename V>
struct VecAt {
    gptr<const V> v;
    __device__ __forceinline__ V operator()(int j) const { return v[j]; }
};

// Row sum of SELL row `lane` of slice `s` against operand f (wave-uniform slice).
template <typename T, typename F>
__device__ __forceinline__ T sell_row(const GDesc &D, gptr<const T> val, const F &v, int s, int lane, int len) {
    const int base = D.sl_ptr[s];
    const int L = (D.sl_ptr[s + 1] - base) >> 6;
    T acc = 0;
    for (int k = 0; k < L; k += kSellU) {
        int c[kSellU];
        T a[kSellU];
        T xv[kSellU];
#pragma unroll
        for (int u = 0; u < kSellU; ++u) {
            const int idx = base + min(k + u, L - 1) * 64 + lane;
            c[u] = __builtin_nontemporal_load(&D.sl_col[idx]);
            a[u] = __builtin_nontemporal_load(&val[idx]);
        }
#pragma unroll
        for (int u = 0; u < kSellU; ++u) xv[u] = v(c[u]);
#pragma unroll
        for (int u = 0; u < kSellU; ++u)
            if (k + u < len) acc = badd(acc, bmul(a[u], xv[u]));
    }
    return acc;
}

// The same row sum over coded columns: column = min(r + sdict[code], n-1)
// (padding slots hold kPadCode, offset 0 in sdict, and are not summed, so
// no row length is loaded; rows past n read x[n-1] and are not stored).
// Word w of a lane (slots 4w..4w+3) sits at sl_wptr[s] + 64 w + lane; the
// values keep the slice width L.  Two words (8 slots) per step, the slots
// past L skipped (wave-uniform).
template <typename T, typename F>
__device__ __forceinline__ T sell_row_coded(const GDesc &D, gptr<const T> val, const F &v, int s, int lane, int r,
                                            const int *sdict) {
    const int base = D.sl_ptr[s];
    const int L = (D.sl_ptr[s + 1] - base) >> 6;
    const int wb = D.sl_wptr[s];
    T acc = 0;
    for (int w = 0; 4 * w < L; w += 2) {
        unsigned cw[2];
        T a[8];
        T xv[8];
#pragma unroll
        for (int h = 0; h < 2; ++h) cw[h] = 4 * (w + h) < L ? __builtin_nontemporal_load(&D.sl_code[wb + (w + h) * 64 + lane]) : 0u;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = 4 * w + u;
            a[u] = 0;
            xv[u] = 0;
            if (k < L) {   // wave-uniform
                a[u] = __builtin_nontemporal_load(&val[base + k * 64 + lane]);
                xv[u] = v(min(r + sdict[(cw[u >> 2] >> (8 * (u & 3))) & 255u], D.n - 1));
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (4 * w + u < L && ((cw[u >> 2] >> (8 * (u & 3))) & 255u) != kPadCode) acc = badd(acc, bmul(a[u], xv[u]));
    }
    return acc;
}

__device__ __forceinline__ bool is_inner(int ph) { return ph == P_START || ph == P_GO; }
__device__ __forceinline__ bool is_outer(int ph) { return ph == P_FIRST || ph == P_CORRECT; }
__device__ __forceinline__ bool works(int ph) { return ph >= P_FIRST && ph <= P_GO; }

// The vector step of an inner phase PH over rows [r0, r1) of a tile (rv:
// the end of the rows taken 4 at a time with 16-byte accesses): P_SETUP
// r = fl(r64); P_START p = u, s = w; P_GO p = u + b p, s = w + b s; then
// c += a p, r -= a s (not in P_SETUP) and u = invd r, with this thread's
// parts of r.u and r.r.
template <int PH>
__device__ __forceinline__ void vec_rows(const GDesc &D, int r0, int r1, int rv, float a, float b, int tid, double &ru,
                                         double &rr) {
    // the current u = invd r is formed again from r (the same bits as the
    // stored copy, which only the SpMV reads)
    auto one = [&](float &ri, float &si, float &pi, float &ci, float &un, float wi, float di, double r64i) {
        if (PH == P_SETUP) {
            ri = static_cast<float>(r64i);
        } else {
            const float ui = bmul(di, ri);
            pi = PH == P_START ? ui : ui + b * pi;
            si = PH == P_START ? wi : wi + b * si;
            ci = ci + a * pi;
            ri = ri - a * si;
        }
        un = bmul(di, ri);
        ru += static_cast<double>(ri) * static_cast<double>(un);
        rr += static_cast<double>(ri) * static_cast<double>(ri);
    };
    for (int i = r0 + 4 * tid; i < rv; i += 4 * kBlock) {
        bf4 r4 = bf4(0.0f), s4 = bf4(0.0f), p4 = bf4(0.0f), c4 = bf4(0.0f), w4 = bf4(0.0f), u4;
        bd2 ra = bd2(0.0), rb = bd2(0.0);
        if (PH == P_SETUP) {
            ra = *reinterpret_cast<gptr<const bd2>>(&D.r64[i]);
            rb = *reinterpret_cast<gptr<const bd2>>(&D.r64[i + 2]);
        } else {
            r4 = *reinterpret_cast<gptr<const bf4>>(&D.r[i]);
            w4 = *reinterpret_cast<gptr<const bf4>>(&D.w[i]);
            c4 = *reinterpret_cast<gptr<const bf4>>(&D.c[i]);
            if (PH == P_GO) {
                s4 = *reinterpret_cast<gptr<const bf4>>(&D.s[i]);
                p4 = *reinterpret_cast<gptr<const bf4>>(&D.p[i]);
            }
        }
        const bf4 d4 = *reinterpret_cast<gptr<const bf4>>(&D.invd[i]);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float ri = r4[k], si = s4[k], pi = p4[k], ci = c4[k], un;
            one(ri, si, pi, ci, un, w4[k], d4[k], k < 2 ? ra[k] : rb[k - 2]);
            r4[k] = ri;
            s4[k] = si;
            p4[k] = pi;
            c4[k] = ci;
            u4[k] = un;
        }
        *reinterpret_cast<gptr<bf4>>(&D.r[i]) = r4;
        *reinterpret_cast<gptr<bf4>>(&D.u[i]) = u4;
        if (PH != P_SETUP) {
            *reinterpret_cast<gptr<bf4>>(&D.s[i]) = s4;
            *reinterpret_cast<gptr<bf4>>(&D.p[i]) = p4;
            *reinterpret_cast<gptr<bf4>>(&D.c[i]) = c4;
        }
    }
    for (int i = rv + tid; i < r1; i += kBlock) {
        float ri = PH == P_SETUP ? 0.0f : D.r[i], si = PH == P_GO ? D.s[i] : 0.0f, pi = PH == P_GO ? D.p[i] : 0.0f;
        float ci = PH == P_SETUP ? 0.0f : D.c[i], un;
        one(ri, si, pi, ci, un, PH == P_SETUP ? 0.0f : D.w[i], D.invd[i], PH == P_SETUP ? D.r64[i] : 0.0);
        D.r[i] = ri;
        D.u[i] = un;
        if (PH != P_SETUP) {
            D.s[i] = si;
            D.p[i] = pi;
            D.c[i] = ci;
        }
    }
}

// One tile's share of K1 for a group of kBlock threads (tid 0..kBlock-1):
// the inner step's vector updates, the start of an inner solve, or the
// correction of x; this thread's parts of the tile's r.u and r.r.
__device__ __forceinline__ void vec_tile(const GDesc &D, const Tile &T, const SysState &S, int tid, double &ru,
                                         double &rr) {
    const int ph = S.phase;
    const int r0 = T.s0 * 64, r1 = min(D.n, (T.s0 + T.ns) * 64);
    // 4 rows per thread and step with 16-byte accesses (a tile starts at a
    // multiple of 2048 rows and every array at a 256-byte boundary), the
    // rows past the last multiple of 4 one by one
    const int rv = r0 + ((r1 - r0) & ~3);
    if (is_outer(ph)) {
        for (int i = r0 + 4 * tid; i < rv; i += 4 * kBlock) {
            if (ph == P_CORRECT) {
                const bf4 c4 = *reinterpret_cast<gptr<const bf4>>(&D.c[i]);
                bd2 xa = *reinterpret_cast<gptr<const bd2>>(&D.x[i]);
                bd2 xb = *reinterpret_cast<gptr<const bd2>>(&D.x[i + 2]);
                xa.x += static_cast<double>(c4.x);
                xa.y += static_cast<double>(c4.y);
                xb.x += static_cast<double>(c4.z);
                xb.y += static_cast<double>(c4.w);
                *reinterpret_cast<gptr<bd2>>(&D.x[i]) = xa;
                *reinterpret_cast<gptr<bd2>>(&D.x[i + 2]) = xb;
            }
            *reinterpret_cast<gptr<bf4>>(&D.c[i]) = bf4(0.0f);
        }
        for (int i = rv + tid; i < r1; i += kBlock) {
            if (ph == P_CORRECT) D.x[i] += static_cast<double>(D.c[i]);
            D.c[i] = 0.0f;
        }
        return;
    }
    const float a = static_cast<float>(S.alpha), b = static_cast<float>(S.beta);
    if (ph == P_SETUP)
        vec_rows<P_SETUP>(D, r0, r1, rv, a, b, tid, ru, rr);
    else if (ph == P_START)
        vec_rows<P_START>(D, r0, r1, rv, a, b, tid, ru, rr);
    else
        vec_rows<P_GO>(D, r0, r1, rv, a, b, tid, ru, rr);
}

// w = A32 u over a tile of a dictionary-coded system, with this thread's part
// of w.u: the rows of sell_row_coded<float> in the same order, but a wave
// (which walks T.ns / 4 slices of the tile) loads its next slice's header,
// code words and values right after its current gathers, so their latency
// overlaps those gathers.
__device__ __forceinline__ double spmv_tile_coded32(const GDesc &D, const Tile &T, int wave, int lane,
                                                    const int *sdict) {
    const VecAt<float> f{gptr<const float>(D.u)};
    const gptr<const float> val = D.v32;
    constexpr int kStride = kBlock / 64;
    const int se = T.s0 + T.ns;
    int nb = 0, nL = 0, nwb = 0;
    unsigned ncw[2];
    float na[8];
    auto fetch = [&](int sn) {   // word 0..1 and slots 0..7 of slice sn
        nb = D.sl_ptr[sn];
        nL = (D.sl_ptr[sn + 1] - nb) >> 6;
        nwb = D.sl_wptr[sn];
#pragma unroll
        for (int h = 0; h < 2; ++h)
            ncw[h] = 4 * h < nL ? __builtin_nontemporal_load(&D.sl_code[nwb + h * 64 + lane]) : 0u;
#pragma unroll
        for (int u = 0; u < 8; ++u) na[u] = u < nL ? __builtin_nontemporal_load(&val[nb + u * 64 + lane]) : 0.0f;
    };
    double dsum = 0;
    int s = T.s0 + wave;
    if (s < se) fetch(s);
    for (; s < se; s += kStride) {
        const int base = nb, L = nL, wb = nwb;
        const int r = s * 64 + lane;
        float acc = 0;
        for (int w = 0; 4 * w < L; w += 2) {
            unsigned cw[2];
            float a[8], xv[8];
            if (w == 0) {
#pragma unroll
                for (int h = 0; h < 2; ++h) cw[h] = ncw[h];
#pragma unroll
                for (int u = 0; u < 8; ++u) a[u] = na[u];
            } else {
#pragma unroll
                for (int h = 0; h < 2; ++h)
                    cw[h] = 4 * (w + h) < L ? __builtin_nontemporal_load(&D.sl_code[wb + (w + h) * 64 + lane]) : 0u;
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    a[u] = 4 * w + u < L ? __builtin_nontemporal_load(&val[base + (4 * w + u) * 64 + lane]) : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)   // wave-uniform test
                xv[u] = 4 * w + u < L ? f(min(r + sdict[(cw[u >> 2] >> (8 * (u & 3))) & 255u], D.n - 1)) : 0.0f;
            if (4 * (w + 2) >= L && s + kStride < se) fetch(s + kStride);   // after this step's gathers
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (4 * w + u < L && ((cw[u >> 2] >> (8 * (u & 3))) & 255u) != kPadCode)
                    acc = badd(acc, bmul(a[u], xv[u]));
        }
        if (L == 0 && s + kStride < se) fetch(s + kStride);   // an empty slice ran no step
        if (r < D.n) {
            D.w[r] = acc;
            dsum += static_cast<double>(acc) * static_cast<double>(f(r));
        }
    }
    return dsum;
}

// One tile's share of K2 for a group of kBlock threads (wave 0..3 of the
// group, wave-uniform): w = A32 u with this thread's part of w.u, or the
// FP64 true residual r64 = b - A64 x with its part of r64.r64.
__device__ __forceinline__ double spmv_tile(const GDesc &D, const Tile &T, int ph, int wave, int lane,
                                            const int *sdict) {
    const bool coded = D.ndict > 0;
    const bool outer = is_outer(ph);
    double dsum = 0;
    if (coded && !outer) return spmv_tile_coded32(D, T, wave, lane, sdict);
    for (int s = T.s0 + wave; s < T.s0 + T.ns; s += kBlock / 64) {
        const int r = s * 64 + lane;
        const int len = !coded && r < D.n ? D.sl_len[r] : 0;   // coded: padding slots carry kPadCode
        if (!outer) {
            const VecAt<float> f{gptr<const float>(D.u)};
            const float acc = coded ? sell_row_coded<float>(D, D.v32, f, s, lane, r, sdict)
                                    : sell_row<float>(D, D.v32, f, s, lane, len);
            if (r < D.n) {
                D.w[r] = acc;
                dsum += static_cast<double>(acc) * static_cast<double>(f(r));
            }
        } else {
            const VecAt<double> f{gptr<const double>(D.x)};
            const double acc = coded ? sell_row_coded<double>(D, D.v64, f, s, lane, r, sdict)
                                     : sell_row<double>(D, D.v64, f, s, lane, len);
            if (r < D.n) {
                const double rr = D.b[r] - acc;
                D.r64[r] = rr;
                dsum += rr * rr;
            }
        }
    }
    return dsum;
}

// K1: every active system's tiles: the vector step; partials r.u (slot 0)
// and r.r (slot 2) of the inner phases.
__global__ __launch_bounds__(kBlock) void kb_vec(const SysDesc *__restrict__ sd, const SysState *__restrict__ st,
                                                 const Tile *__restrict__ tiles, const int *__restrict__ ntiles) {
    __shared__ double sh[kBlock / 64];
    const int nt = *ntiles;
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        const Tile T = tiles[t];
        const SysState S = st[T.sys];
        if (!works(S.phase)) continue;   // uniform over the workgroup
        const GDesc D = gview(sd[T.sys]);
        double ru = 0, rr = 0;
        vec_tile(D, T, S, threadIdx.x, ru, rr);
        if (is_outer(S.phase)) continue;
        const double tru = block_sum(ru, sh);
        __syncthreads();
        const double trr = block_sum(rr, sh);
        if (threadIdx.x == 0) {
            D.part[kParts * T.lt] = tru;
            D.part[kParts * T.lt + 2] = trr;
        }
        __syncthreads();
    }
}

// K2: every active system's tiles: the SpMV; partial w.u (slot 1) or
// r64.r64 (slot 2).
__global__ __launch_bounds__(kBlock) void kb_spmv(const SysDesc *__restrict__ sd, const SysState *__restrict__ st,
                                                  const Tile *__restrict__ tiles, const int *__restrict__ ntiles) {
    __shared__ double sh[kBlock / 64];
    __shared__ int sdict[256];
    const int nt = *ntiles;
    // the wave index in a scalar register: slice quantities are then scalar
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        const Tile T = tiles[t];
        const int ph = st[T.sys].phase;
        if (!works(ph)) continue;   // uniform over the workgroup
        const GDesc D = gview(sd[T.sys]);
        if (D.ndict > 0) {   // the system's offset table (the previous tile's readers passed the barrier below)
            for (int i = threadIdx.x; i < 256; i += kBlock) sdict[i] = i < D.ndict ? D.dict[i] : 0;
            __syncthreads();
        }
        const double tot = block_sum(spmv_tile(D, T, ph, wave, lane, sdict), sh);
        if (threadIdx.x == 0) D.part[kParts * T.lt + (is_outer(ph) ? 2 : 1)] = tot;
        __syncthreads();   // sh and sdict are reused by the next tile
    }
}

// The decision after K1 and K2 from the sums of their partials (g = r.u,
// d = w.u, rho = r.r, or the true residual's r64.r64).
__device__ __forceinline__ void decide(SysState &S, double g, double d, double rho, const KParams &prm) {
    const int ph = S.phase;
    if (is_outer(ph)) {
        const double res = std::sqrt(rho);
        if (ph == P_FIRST) {
            S.res0 = res;
            S.tol = fmax(prm.atol, prm.rtol * res);
        }
        S.res = res;
        if (res <= S.tol) {
            S.status = res <= prm.atol ? HFH_SUCCESS_ABS : HFH_SUCCESS_REL;
            S.phase = P_DONE;
        } else if (!isfinite(res)) {
            S.status = HFH_FAIL_DIV;
            S.phase = P_DONE;
        } else if (S.outer >= prm.outer_max) {
            S.status = HFH_FAIL_MAXITS;
            S.phase = P_DONE;
        } else {
            ++S.outer;
            S.phase = P_SETUP;
        }
        return;
    }
    const double rn = std::sqrt(rho);
    if (ph == P_SETUP) {
        S.rr0 = rn;
        S.inner = 0;
        if (rn > 0 && isfinite(rn) && d > 0 && isfinite(d) && isfinite(g)) {
            S.alpha = g / d;
            S.beta = 0;
            S.gamma = g;
            S.phase = P_START;
        } else {
            S.phase = P_CORRECT;   // c = 0: the true residual of x again
        }
        return;
    }
    ++S.inner;
    ++S.inner_total;
    if (rn <= prm.inner_rtol * S.rr0 || S.inner >= prm.inner_max || !(g != 0) || !isfinite(g)) {
        S.phase = P_CORRECT;
        return;
    }
    const double beta = g / S.gamma;
    const double den = d - beta * g / S.alpha;
    if (!(den > 0) || !isfinite(den)) {
        S.phase = P_CORRECT;   // breakdown: end this inner solve with the correction so far
        return;
    }
    S.beta = beta;
    S.alpha = g / den;
    S.gamma = g;
    S.phase = P_GO;
}

// K3: per system, after K1 and K2.
// The slot's phase and its descriptor's fields are loaded together, then
// all its partials in one pass.  Each quantity is summed in a fixed order:
// threads strided over the tiles in ascending order, the wave trees, then
// 0 + the four wave totals in order (kb_small's group 0 sums the same way).
__global__ __launch_bounds__(kBlock) void kb_fin(const SysDesc *__restrict__ sd, SysState *__restrict__ st,
                                                 int nslots, KParams prm) {
    __shared__ double sh[kParts][kBlock / 64];
    const int s = blockIdx.x;
    if (s >= nslots) return;
    const int ph = st[s].phase;
    const int small = sd[s].small, nt = sd[s].ntiles;
    const gptr<const double> part = (gptr<const double>)sd[s].part;
    if (!works(ph) || small) return;   // idle or finished, or kb_small's
    const int q0 = is_outer(ph) ? 2 : 0;
    double v[kParts] = {0, 0, 0};
    for (int i = threadIdx.x; i < nt; i += kBlock)
#pragma unroll
        for (int q = 0; q < kParts; ++q)
            if (q >= q0) v[q] += part[kParts * i + q];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < kParts; ++q) {
        const double t = wave_sum(v[q]);
        if (lane == 0) sh[q][w] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot[kParts];
#pragma unroll
        for (int q = 0; q < kParts; ++q) {
            tot[q] = 0;
            for (int i = 0; i < kBlock / 64; ++i) tot[q] += sh[q][i];
        }
        decide(st[s], tot[0], tot[1], tot[2], prm);
    }
}

// Sum of a group of kBlock threads (the 4 waves from `gw`), with block_sum's
// order: wave sums, then 0 + w0 + w1 + w2 + w3; valid in every thread of the
// group.  Every thread of the workgroup calls it (it holds a barrier).
__device__ __forceinline__ double group_sum(double v, double *gsh /* kBlock / 64 */, int gwave, int lane) {
    v = wave_sum(v);
    if (lane == 0) gsh[gwave] = v;
    __syncthreads();
    double s = 0;
    for (int i = 0; i < kBlock / 64; ++i) s += gsh[i];
    return s;
}

// Systems of at most kSmallTiles tiles: one workgroup per system runs all
// `iters` steps of a replay, K1 to K3 with barriers in place of launches.
// Group g of kBlock threads does what the tile kernels' workgroup does for
// the system's tile g, and group 0 what the slot kernel does, in the same
// order of operations, so the results are those of the tile path bit for bit.
constexpr int kSmallTiles = 4;
constexpr int kSmallBlock = kSmallTiles * kBlock;
// steps of kb_small per replay, in chunks: a small system's steps are short,
// and the replay lasts as long as the tile launches anyway
constexpr int kSmallChunks = 4;

__global__ __launch_bounds__(kSmallBlock) void kb_small(const SysDesc *__restrict__ sd, SysState *__restrict__ st,
                                                        const int *__restrict__ small, int nsmall, int iters,
                                                        KParams prm) {
    __shared__ double gsh[kSmallTiles][kBlock / 64];
    __shared__ int sdict[256];
    __shared__ SysState S;
    __shared__ double tot[kParts];
    const int g = __builtin_amdgcn_readfirstlane(threadIdx.x / kBlock);
    const int tid = threadIdx.x % kBlock;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    for (int k = blockIdx.x; k < nsmall; k += gridDim.x) {
        const int slot = small[k];
        const GDesc D = gview(sd[slot]);
        for (int i = threadIdx.x; i < 256; i += kSmallBlock) sdict[i] = i < D.ndict ? D.dict[i] : 0;
        if (threadIdx.x == 0) S = st[slot];
        __syncthreads();
        const bool mine = g < D.ntiles;   // this group's tile exists
        const Tile T{slot, g * kTileSlices, min(kTileSlices, D.nslice - g * kTileSlices), g};
        for (int it = 0; it < iters; ++it) {
            const SysState Sk = S;
            const int ph = Sk.phase;
            if (!works(ph)) break;   // uniform
            // K1
            double ru = 0, rr = 0;
            if (mine) vec_tile(D, T, Sk, tid, ru, rr);
            if (!is_outer(ph)) {
                const double tru = group_sum(ru, gsh[g], wave, lane);
                __syncthreads();
                const double trr = group_sum(rr, gsh[g], wave, lane);
                if (tid == 0 && mine) {
                    D.part[kParts * g] = tru;
                    D.part[kParts * g + 2] = trr;
                }
            }
            __syncthreads();   // K2 gathers what K1 wrote across the groups
            // K2
            const double tsp = group_sum(mine ? spmv_tile(D, T, ph, wave, lane, sdict) : 0.0, gsh[g], wave, lane);
            if (tid == 0 && mine) D.part[kParts * g + (is_outer(ph) ? 2 : 1)] = tsp;
            __syncthreads();
            // K3: the slot kernel's sums, by group 0
            for (int q = is_outer(ph) ? 2 : 0; q < kParts; ++q) {
                double v = 0;
                if (g == 0)
                    for (int i = tid; i < D.ntiles; i += kBlock) v += D.part[kParts * i + q];
                const double t = group_sum(v, gsh[g], wave, lane);
                if (threadIdx.x == 0) tot[q] = t;
                __syncthreads();
            }
            if (threadIdx.x == 0) decide(S, is_outer(ph) ? 0.0 : tot[0], is_outer(ph) ? 0.0 : tot[1], tot[2], prm);
            __syncthreads();   // S, and the next K1 reads what this K2 wrote
        }
        if (threadIdx.x == 0) st[slot] = S;
        __syncthreads();
    }
}

// Writes one slot's descriptor (when `desc`) and state from the kernel
// arguments: stream-ordered behind the running iterations, with no host
// staging buffer to keep alive.
__global__ void kb_set_slot(SysDesc *__restrict__ sd, SysState *__restrict__ st, int slot, int desc, SysDesc d,
                            SysState v) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (desc) sd[slot] = d;
    st[slot] = v;
}

// Device assembly of the cell-centred FV diffusion matrix (5-pt when nz == 1,
// else 7-pt) in SELL-64 form of uniform width W from the cell coefficients:
// the same entries, in the same order and with the same arithmetic as the
// host generator (harmonic-mean faces, 2 k on Dirichlet faces), the columns
// as one-byte codes into the offset table (written here too).  One thread
// per row slot; rows beyond n (last slice) are padding.
__global__ __launch_bounds__(kBlock) void kb_fv_assemble(int nx, int ny, int nz, int n, int nslice, int W,
                                                         const double *__restrict__ kc, int *__restrict__ sl_ptr,
                                                         int *__restrict__ sl_len, unsigned *__restrict__ sl_code,
                                                         int *__restrict__ sl_wptr, int *__restrict__ dict,
                                                         double *__restrict__ v64, float *__restrict__ v32,
                                                         float *__restrict__ invd) {
#pragma clang fp contract(off)
    const bool in3 = nz > 1;
    const long long nxy = static_cast<long long>(nx) * ny;
    // offset table, in the order of a row's entries: -nxy, -nx, -1, 0, 1, nx, nxy
    // (2D: without +-nxy); code0 = the code of offset 0 (the diagonal);
    // padding slots and rows past n: kPadCode
    const int nd = in3 ? 7 : 5, code0 = in3 ? 3 : 2;
    const int WW = (W + 3) / 4;   // code words per lane and slice
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r <= nslice && r >= 0) {
        // slice and word pointers (nslice+1 of each) from the first threads
        sl_ptr[r] = r * 64 * W;
        sl_wptr[r] = r * 64 * WW;
    }
    if (r < nd) {
        const long long tab3[7] = {-nxy, -nx, -1, 0, 1, nx, nxy};
        const long long tab2[5] = {-static_cast<long long>(nx), -1, 0, 1, nx};
        dict[r] = static_cast<int>(in3 ? tab3[r] : tab2[r]);
    }
    if (r >= nslice * 64) return;
    const int s = r >> 6, lane = r & 63;
    const int base = s * 64 * W, wbase = s * 64 * WW;
    unsigned words[2] = {0u, 0u};   // WW <= 2
    auto code_at = [&](int e, unsigned c) { words[e >> 2] |= c << (8 * (e & 3)); };
    if (r >= n) {
        for (int e = 0; e < W; ++e) {
            v64[base + e * 64 + lane] = 0.0;
            v32[base + e * 64 + lane] = 0.0f;
        }
        for (int e = 0; e < 4 * WW; ++e) code_at(e, kPadCode);
        for (int w = 0; w < WW; ++w) sl_code[wbase + w * 64 + lane] = words[w];
        return;
    }
    const int i = r % nx, j = (r / nx) % ny, k = r / (nx * ny);
    const long long nb[6] = {r - nxy, r - nx, r - 1LL, r + 1LL, r + nx, r + nxy};
    const bool has[6] = {in3 && k > 0, j > 0, i > 0, i < nx - 1, j < ny - 1, in3 && k < nz - 1};
    const bool axis[6] = {in3, true, true, true, true, in3};
    // code of neighbour q (the offset table skips +-nxy in 2D)
    const unsigned qcode[6] = {0u, in3 ? 1u : 0u, in3 ? 2u : 1u, in3 ? 4u : 3u, in3 ? 5u : 4u, 6u};
    const double ks = kc[r];
    double off[6] = {0, 0, 0, 0, 0, 0};
    double diag = 0;
    for (int q = 0; q < 6; ++q) {
        if (!axis[q]) continue;
        if (has[q]) {
            const double kn = kc[nb[q]];
            const double t = 2.0 * ks * kn / (ks + kn);
            off[q] = -t;
            diag += t;
        } else {
            diag += 2.0 * ks;
        }
    }
    int e = 0;
    auto put = [&](unsigned code, double v) {
        const int idx = base + e * 64 + lane;
        code_at(e, code);
        v64[idx] = v;
        v32[idx] = static_cast<float>(v);
        ++e;
    };
    for (int q = 0; q < 3; ++q)
        if (has[q]) put(qcode[q], off[q]);
    put(static_cast<unsigned>(code0), diag);
    for (int q = 3; q < 6; ++q)
        if (has[q]) put(qcode[q], off[q]);
    sl_len[r] = e;
    for (int f = e; f < W; ++f) {
        const int idx = base + f * 64 + lane;
        v64[idx] = 0.0;
        v32[idx] = 0.0f;
    }
    for (int f = e; f < 4 * WW; ++f) code_at(f, kPadCode);
    for (int w = 0; w < WW; ++w) sl_code[wbase + w * 64 + lane] = words[w];
    invd[r] = 1.0f / static_cast<float>(diag);
}

// host-side first-fit allocator over [0, cap) with coalescing
struct Arena {
    size_t cap = 0, used = 0;
    std::map<size_t, size_t> freel;   // offset -> size
    void init(size_t c) {
        cap = c;
        used = 0;
        freel.clear();
        if (c) freel[0] = c;
    }
    bool alloc(size_t bytes, size_t *off) {
        for (auto it = freel.begin(); it != freel.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes;
            const size_t at = it->first + bytes;
            freel.erase(it);
            if (rest) freel[at] = rest;
            used += bytes;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        used -= bytes;
        auto it = freel.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != freel.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            freel.erase(nx);
        }
        if (it != freel.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                freel.erase(it);
            }
        }
    }
};

size_t up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

// byte offsets of a system's sections inside its region
// words > 0: coded columns (`words` code words, a word pointer per slice and
// the offset table) instead of E column indices
struct Layout {
    size_t ptr, len, col, wptr, dict, v64, v32, invd, b, x, upload, r64, r, w, s, u, c, p, part, total;
    long long words;
    Layout(int n, int nslice, long long E, int ntiles, long long words_ = 0) : words(words_) {
        size_t o = 0;
        const size_t vf = up(sizeof(float) * static_cast<size_t>(n));
        ptr = o; o += up(sizeof(int) * (nslice + 1));
        len = o; o += up(sizeof(int) * static_cast<size_t>(n));
        col = o; o += up(words ? sizeof(unsigned) * static_cast<size_t>(words) : sizeof(int) * static_cast<size_t>(E));
        wptr = o; o += words ? up(sizeof(int) * (nslice + 1)) : 0;
        dict = o; o += words ? up(sizeof(int) * 256) : 0;
        v64 = o; o += up(sizeof(double) * static_cast<size_t>(E));
        v32 = o; o += up(sizeof(float) * static_cast<size_t>(E));
        invd = o; o += vf;
        b = o; o += up(sizeof(double) * static_cast<size_t>(n));
        x = o; o += up(sizeof(double) * static_cast<size_t>(n));
        upload = o;
        r64 = o; o += up(sizeof(double) * static_cast<size_t>(n));
        r = o; o += vf;
        w = o; o += vf;
        s = o; o += vf;
        u = o; o += vf;
        c = o; o += vf;
        p = o; o += vf;
        part = o; o += up(sizeof(double) * kParts * static_cast<size_t>(ntiles));
        total = o;
    }
};

int tiles_of(int nslice) { return (nslice + kTileSlices - 1) / kTileSlices; }

// SELL-64 shape of a host CSR matrix: slice widths and padded entry count
long long sell_entries(int n, const int *row, std::vector<int> *ptr) {
    const int nsl = (n + 63) / 64;
    if (ptr) ptr->assign(nsl + 1, 0);
    long long off = 0;
    for (int s = 0; s < nsl; ++s) {
        int L = 0;
        for (int r = s * 64; r < std::min(n, s * 64 + 64); ++r) L = std::max(L, row[r + 1] - row[r]);
        if (ptr) (*ptr)[s] = static_cast<int>(std::min<long long>(off, 0x7fffffffLL));
        off += 64LL * L;
    }
    if (ptr) (*ptr)[nsl] = static_cast<int>(std::min<long long>(off, 0x7fffffffLL));
    return off;
}

}  // namespace
}  // namespace hfh

using namespace hfh;

struct hfh_batch {
    hfh_ctx *ctx = nullptr;
    hfh_batch_params prm{};
    hipStream_t cs = nullptr;        // uploads, device assembly and solution copies
    hipStream_t ss = nullptr;        // kb_small, beside the tile launches on the context stream
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int small_tiles = 0;             // systems of at most this many tiles run in kb_small
    int *d_small = nullptr;          // slots of the small systems (kb_small's list)
    int *h_small[2] = {nullptr, nullptr};
    int nsmall_now = 0, ntiles_now = 0;   // the tables the next replay reads
    char *arena_d = nullptr;
    Arena arena;
    SysDesc *d_sys = nullptr;
    SysState *d_st = nullptr;
    Tile *d_tiles = nullptr;
    int *d_ntiles = nullptr;
    long long tile_cap = 0;
    struct Slot {
        int id = -1;          // request id, -1 when free
        size_t off = 0, bytes = 0;
        int n = 0, nslice = 0, ntiles = 0;
        bool small = false;   // solved by kb_small
        bool reported = false;
        double *x = nullptr;  // device solution vector
        SysState final_state{};
        int seen_inner = 0;   // inner iterations counted into row_iters so far
    };
    std::vector<Slot> slots;
    // state readback, double buffered: replay k's copy is read while replay k+1 runs
    SysState *h_st[2] = {nullptr, nullptr};
    hipEvent_t ev_rb[2] = {nullptr, nullptr};
    int rb_next = 0, rb_pending = -1;
    // tile table staging, double buffered
    Tile *h_tiles[2] = {nullptr, nullptr};
    int *h_ntiles[2] = {nullptr, nullptr};
    hipEvent_t ev_tiles[2] = {nullptr, nullptr};
    int tb = 0;
    // upload staging: free again once ev_up has passed
    void *stage = nullptr;
    size_t stage_cap = 0;
    hipEvent_t ev_up = nullptr;
    bool tiles_dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<int> finished;       // ids finished and not yet collected
    std::map<int, int> slot_of;      // id -> slot
    int next_id = 1;
    long long steps = 0;
    // statistics: GPU time of the replays (events around each graph launch,
    // read with its state copy) and inner row-iterations (sum of n x inner
    // iterations) seen in the state copies
    hipEvent_t ev_g0[2] = {nullptr, nullptr}, ev_g1[2] = {nullptr, nullptr};
    double busy_ms = 0;
    long long replays = 0, row_iters = 0;
};

namespace {

int ensure_stage(hfh_batch *B, size_t bytes) {
    if (bytes <= B->stage_cap) return HFH_OK;
    if (B->stage) HFH_HIP(hipHostFree(B->stage));
    B->stage = nullptr;
    B->stage_cap = 0;
    size_t cap = std::max(bytes, static_cast<size_t>(64) << 20);
    HFH_HIP(hipHostMalloc(&B->stage, cap, hipHostMallocDefault));
    B->stage_cap = cap;
    return HFH_OK;
}

int free_slot(const hfh_batch *B) {
    for (size_t s = 0; s < B->slots.size(); ++s)
        if (B->slots[s].id < 0) return static_cast<int>(s);
    return -1;
}

// Reserve a slot and an arena region; fills *slot and *off.
int reserve(hfh_batch *B, size_t bytes, int *slot, size_t *off) {
    HFH_CHECK(bytes <= B->arena.cap, HFH_ERR_INVALID, "hfh_batch: request of %zu bytes exceeds the arena of %zu bytes",
              bytes, B->arena.cap);
    const int s = free_slot(B);
    HFH_CHECK(s >= 0, HFH_ERR_BUDGET, "hfh_batch: all %d slots are busy", static_cast<int>(B->slots.size()));
    HFH_CHECK(B->arena.alloc(bytes, off), HFH_ERR_BUDGET, "hfh_batch: %zu bytes do not fit the HBM budget now "
              "(%zu of %zu in use)", bytes, B->arena.used, B->arena.cap);
    *slot = s;
    return HFH_OK;
}

// Makes an uploaded system live: the main stream waits for the upload, then
// writes the slot's descriptor and initial state; the tile table is rebuilt
// before the next replay.
int activate(hfh_batch *B, int slot, size_t off, const Layout &L, int n, int nslice, size_t bytes, int ndict, int *id) {
    hipStream_t s = B->ctx->stream;
    char *base = B->arena_d + off;
    SysDesc d{};
    d.n = n;
    d.nslice = nslice;
    d.ntiles = tiles_of(nslice);
    d.small = d.ntiles <= B->small_tiles;
    d.sl_ptr = reinterpret_cast<const int *>(base + L.ptr);
    d.sl_len = reinterpret_cast<const int *>(base + L.len);
    if (L.words) {
        d.sl_code = reinterpret_cast<const unsigned *>(base + L.col);
        d.sl_wptr = reinterpret_cast<const int *>(base + L.wptr);
        d.dict = reinterpret_cast<const int *>(base + L.dict);
        d.ndict = ndict;
    } else {
        d.sl_col = reinterpret_cast<const int *>(base + L.col);
    }
    d.v64 = reinterpret_cast<const double *>(base + L.v64);
    d.v32 = reinterpret_cast<const float *>(base + L.v32);
    d.invd = reinterpret_cast<const float *>(base + L.invd);
    d.b = reinterpret_cast<const double *>(base + L.b);
    d.x = reinterpret_cast<double *>(base + L.x);
    d.r64 = reinterpret_cast<double *>(base + L.r64);
    d.r = reinterpret_cast<float *>(base + L.r);
    d.w = reinterpret_cast<float *>(base + L.w);
    d.s = reinterpret_cast<float *>(base + L.s);
    d.u = reinterpret_cast<float *>(base + L.u);
    d.c = reinterpret_cast<float *>(base + L.c);
    d.p = reinterpret_cast<float *>(base + L.p);
    d.part = reinterpret_cast<double *>(base + L.part);
    hfh_batch::Slot &S = B->slots[slot];
    S.id = B->next_id++;
    SysState st{};
    st.phase = P_FIRST;
    st.id = S.id;
    HFH_HIP(hipStreamWaitEvent(s, B->ev_up, 0));
    hipLaunchKernelGGL(kb_set_slot, dim3(1), dim3(64), 0, s, B->d_sys, B->d_st, slot, 1, d, st);
    int rc = launch_check("kb_set_slot");
    if (rc) {
        S = hfh_batch::Slot();
        return rc;
    }
    S.off = off;
    S.bytes = bytes;
    S.n = n;
    S.nslice = nslice;
    S.ntiles = d.ntiles;
    S.small = d.small != 0;
    S.reported = false;
    S.x = d.x;
    B->slot_of[S.id] = slot;
    B->tiles_dirty = true;
    *id = S.id;
    return HFH_OK;
}

int upload_tiles(hfh_batch *B) {
    hipStream_t s = B->ctx->stream;
    const int buf = B->tb;
    // the copy from this buffer two uploads ago must have left it
    HFH_HIP(hipEventSynchronize(B->ev_tiles[buf]));
    Tile *ht = B->h_tiles[buf];
    int *hs = B->h_small[buf];
    long long nt = 0;
    int ns = 0;
    for (size_t k = 0; k < B->slots.size(); ++k) {
        const hfh_batch::Slot &S = B->slots[k];
        if (S.id < 0) continue;
        if (S.small) {
            hs[ns++] = static_cast<int>(k);
            continue;
        }
        for (int t = 0; t < S.ntiles; ++t) {
            HFH_CHECK(nt < B->tile_cap, HFH_ERR_INVALID, "hfh_batch: tile table full");
            const int s0 = t * kTileSlices;
            ht[nt++] = Tile{static_cast<int>(k), s0, std::min(kTileSlices, S.nslice - s0), t};
        }
    }
    *B->h_ntiles[buf] = static_cast<int>(nt);
    if (nt) HFH_HIP(hipMemcpyAsync(B->d_tiles, ht, sizeof(Tile) * nt, hipMemcpyHostToDevice, s));
    if (ns) HFH_HIP(hipMemcpyAsync(B->d_small, hs, sizeof(int) * ns, hipMemcpyHostToDevice, s));
    B->nsmall_now = ns;
    B->ntiles_now = static_cast<int>(nt);
    HFH_HIP(hipMemcpyAsync(B->d_ntiles, B->h_ntiles[buf], sizeof(int), hipMemcpyHostToDevice, s));
    HFH_HIP(hipEventRecord(B->ev_tiles[buf], s));
    B->tb ^= 1;
    B->tiles_dirty = false;
    return HFH_OK;
}

int enqueue_iterations(hfh_batch *B, int iters) {
    hipStream_t s = B->ctx->stream;
    const int nslots = static_cast<int>(B->slots.size());
    const KParams kp{B->prm.rtol, B->prm.atol, B->prm.inner_rtol, B->prm.outer_maxits, B->prm.inner_maxits};
    for (int it = 0; it < iters; ++it) {
        hipLaunchKernelGGL(kb_vec, dim3(kTileGrid), dim3(kBlock), 0, s, B->d_sys, B->d_st, B->d_tiles, B->d_ntiles);
        hipLaunchKernelGGL(kb_spmv, dim3(kTileGrid), dim3(kBlock), 0, s, B->d_sys, B->d_st, B->d_tiles, B->d_ntiles);
        hipLaunchKernelGGL(kb_fin, dim3(nslots), dim3(kBlock), 0, s, B->d_sys, B->d_st, nslots, kp);
    }
    return launch_check("hfh_batch step");
}

// Marks the systems a state readback shows finished (the id guards against a
// slot that was collected and reused after the copy was taken).
int take_readback(hfh_batch *B, int buf) {
    HFH_HIP(hipEventSynchronize(B->ev_rb[buf]));
    float ms = 0;
    if (hipEventElapsedTime(&ms, B->ev_g0[buf], B->ev_g1[buf]) == hipSuccess) B->busy_ms += ms;
    ++B->replays;
    const SysState *h = B->h_st[buf];
    int fin = 0;
    for (size_t k = 0; k < B->slots.size(); ++k) {
        hfh_batch::Slot &S = B->slots[k];
        if (S.id >= 0 && h[k].id == S.id && h[k].inner_total > S.seen_inner) {
            B->row_iters += static_cast<long long>(h[k].inner_total - S.seen_inner) * S.n;
            S.seen_inner = h[k].inner_total;
        }
        if (S.id < 0 || S.reported || h[k].id != S.id || h[k].phase != P_DONE) continue;
        S.reported = true;
        S.final_state = h[k];
        B->finished.push_back(S.id);
        ++fin;
    }
    return fin;
}

void release_all(hfh_batch *B) {
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    if (B->arena_d) (void)hipFree(B->arena_d);
    if (B->d_sys) (void)hipFree(B->d_sys);
    if (B->d_st) (void)hipFree(B->d_st);
    if (B->d_tiles) (void)hipFree(B->d_tiles);
    if (B->d_ntiles) (void)hipFree(B->d_ntiles);
    if (B->d_small) (void)hipFree(B->d_small);
    for (int b = 0; b < 2; ++b) {
        if (B->h_small[b]) (void)hipHostFree(B->h_small[b]);
        if (B->h_st[b]) (void)hipHostFree(B->h_st[b]);
        if (B->h_tiles[b]) (void)hipHostFree(B->h_tiles[b]);
        if (B->h_ntiles[b]) (void)hipHostFree(B->h_ntiles[b]);
        if (B->ev_rb[b]) (void)hipEventDestroy(B->ev_rb[b]);
        if (B->ev_g0[b]) (void)hipEventDestroy(B->ev_g0[b]);
        if (B->ev_g1[b]) (void)hipEventDestroy(B->ev_g1[b]);
        if (B->ev_tiles[b]) (void)hipEventDestroy(B->ev_tiles[b]);
    }
    if (B->ev_up) (void)hipEventDestroy(B->ev_up);
    if (B->ev_fork) (void)hipEventDestroy(B->ev_fork);
    if (B->ev_join) (void)hipEventDestroy(B->ev_join);
    if (B->stage) (void)hipHostFree(B->stage);
    if (B->cs) (void)hipStreamDestroy(B->cs);
    if (B->ss) (void)hipStreamDestroy(B->ss);
}

}  // namespace

extern "C" int hfh_batch_create(hfh_ctx *ctx, const hfh_batch_params *p, hfh_batch **out) {
    HFH_CHECK(ctx && p && out, HFH_ERR_INVALID, "hfh_batch_create: null argument");
    HFH_CHECK(p->max_systems > 0 && p->max_systems <= 65536, HFH_ERR_INVALID,
              "hfh_batch_create: max_systems %d not in [1, 65536]", p->max_systems);
    HFH_CHECK(p->hbm_budget >= (1 << 20), HFH_ERR_INVALID, "hfh_batch_create: budget below 1 MiB");
    HFH_CHECK(p->chunk >= 1 && p->chunk <= 256, HFH_ERR_INVALID, "hfh_batch_create: chunk %d not in [1, 256]",
              p->chunk);
    HFH_CHECK(p->outer_maxits > 0 && p->inner_maxits > 0 && p->rtol >= 0 && p->atol >= 0 && p->inner_rtol >= 0,
              HFH_ERR_INVALID, "hfh_batch_create: bad tolerances or iteration limits");
    HFH_CHECK(p->small_tiles >= 0 && p->small_tiles <= kSmallTiles, HFH_ERR_INVALID,
              "hfh_batch_create: small_tiles %d not in [0, %d]", p->small_tiles, kSmallTiles);
    *out = nullptr;
    hfh_batch *B = new hfh_batch();
    B->ctx = ctx;
    B->prm = *p;
    B->slots.resize(p->max_systems);
    B->small_tiles = p->small_tiles;   // 0: every system in the tile launches
    // every row costs at least 68 bytes of the region (one entry), so this
    // bounds the tiles the budget can hold
    B->tile_cap = p->hbm_budget / (68LL * kTileRows) + p->max_systems + 1;
    auto fail = [&](int rc) {
        release_all(B);
        delete B;
        return rc;
    };
    int rc;
    if ((rc = dev_alloc(reinterpret_cast<void **>(&B->arena_d), static_cast<size_t>(p->hbm_budget))) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&B->d_sys), sizeof(SysDesc) * p->max_systems)) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&B->d_st), sizeof(SysState) * p->max_systems)) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&B->d_tiles), sizeof(Tile) * B->tile_cap)) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&B->d_ntiles), sizeof(int))) ||
        (rc = dev_alloc(reinterpret_cast<void **>(&B->d_small), sizeof(int) * p->max_systems)))
        return fail(rc);
    B->arena.init(static_cast<size_t>(p->hbm_budget));
    bool ok = hipStreamCreateWithFlags(&B->cs, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&B->ss, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&B->ev_up, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&B->ev_fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&B->ev_join, hipEventDisableTiming) == hipSuccess;
    for (int b = 0; b < 2 && ok; ++b)
        ok = hipHostMalloc(reinterpret_cast<void **>(&B->h_st[b]), sizeof(SysState) * p->max_systems) == hipSuccess &&
             hipHostMalloc(reinterpret_cast<void **>(&B->h_tiles[b]), sizeof(Tile) * B->tile_cap) == hipSuccess &&
             hipHostMalloc(reinterpret_cast<void **>(&B->h_ntiles[b]), sizeof(int)) == hipSuccess &&
             hipHostMalloc(reinterpret_cast<void **>(&B->h_small[b]), sizeof(int) * p->max_systems) == hipSuccess &&
             hipEventCreateWithFlags(&B->ev_rb[b], hipEventDisableTiming) == hipSuccess &&
             hipEventCreate(&B->ev_g0[b]) == hipSuccess && hipEventCreate(&B->ev_g1[b]) == hipSuccess &&
             hipEventCreateWithFlags(&B->ev_tiles[b], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        set_error("hfh_batch_create: stream, event or pinned host allocation failed");
        return fail(HFH_ERR_NOMEM);
    }
    hipStream_t s = ctx->stream;
    // empty slot table (phase P_IDLE everywhere) and no tiles
    std::memset(B->h_st[0], 0, sizeof(SysState) * p->max_systems);
    *B->h_ntiles[0] = 0;
    if (hipMemcpyAsync(B->d_st, B->h_st[0], sizeof(SysState) * p->max_systems, hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        hipMemcpyAsync(B->d_ntiles, B->h_ntiles[0], sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipEventRecord(B->ev_up, B->cs) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        set_error("hfh_batch_create: initial upload failed");
        return fail(HFH_ERR_HIP);
    }
    // capture `chunk` iterations once; the kernels read the tables at run time
    // (without graphs, hfh_batch_step launches them directly)
    if (graphs_enabled() && hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        set_error("hfh_batch_create: stream capture failed");
        return fail(HFH_ERR_HIP);
    }
    if (graphs_enabled()) {
        const int erc = enqueue_iterations(B, p->chunk);
        const hipError_t e2 = hipStreamEndCapture(s, &B->graph);
        if (erc || e2 != hipSuccess || hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0) != hipSuccess) {
            if (!erc) set_error("hfh_batch_create: graph capture failed");
            return fail(erc ? erc : HFH_ERR_HIP);
        }
    }
    ++ctx->live;
    *out = B;
    return HFH_OK;
}

extern "C" int hfh_batch_destroy(hfh_batch *B) {
    if (!B) return HFH_OK;
    (void)hipStreamSynchronize(B->ctx->stream);
    (void)hipStreamSynchronize(B->cs);
    (void)hipStreamSynchronize(B->ss);
    --B->ctx->live;
    release_all(B);
    delete B;
    return HFH_OK;
}

extern "C" int hfh_batch_bytes_csr(int n, const int *row, long long *bytes) {
    HFH_CHECK(n > 0 && row && bytes, HFH_ERR_INVALID, "hfh_batch_bytes_csr: bad argument");
    const long long E = sell_entries(n, row, nullptr);
    const int nsl = (n + 63) / 64;
    *bytes = static_cast<long long>(Layout(n, nsl, E, tiles_of(nsl)).total);
    return HFH_OK;
}

static int fv_shape(int nx, int ny, int nz, int *n, int *W) {
    HFH_CHECK(nx >= 1 && ny >= 1 && nz >= 1, HFH_ERR_INVALID, "hfh_batch: bad grid %d x %d x %d", nx, ny, nz);
    const long long N = static_cast<long long>(nx) * ny * nz;
    HFH_CHECK(N < (1LL << 30), HFH_ERR_INVALID, "hfh_batch: grid too large");
    *n = static_cast<int>(N);
    *W = nz > 1 ? 7 : 5;
    return HFH_OK;
}

// FV systems store coded columns: 5 or 7 offsets, ceil(W / 4) = 2 code words
// per lane and slice
static long long fv_words(int W, int nsl) { return 64LL * ((W + 3) / 4) * nsl; }

extern "C" int hfh_batch_bytes_fv(int nx, int ny, int nz, long long *bytes) {
    HFH_CHECK(bytes, HFH_ERR_INVALID, "hfh_batch_bytes_fv: null argument");
    int n, W, rc;
    if ((rc = fv_shape(nx, ny, nz, &n, &W))) return rc;
    const int nsl = (n + 63) / 64;
    *bytes = static_cast<long long>(Layout(n, nsl, 64LL * W * nsl, tiles_of(nsl), fv_words(W, nsl)).total);
    return HFH_OK;
}

extern "C" int hfh_batch_submit_csr(hfh_batch *B, int n, const int *row, const int *col, const double *val,
                                    const double *b, const double *x0, int *id) {
    HFH_CHECK(B && n > 0 && row && col && val && b && id, HFH_ERR_INVALID, "hfh_batch_submit_csr: bad argument");
    HFH_CHECK(row[0] == 0, HFH_ERR_INVALID, "hfh_batch_submit_csr: row[0] must be 0");
    for (int r = 0; r < n; ++r) {
        HFH_CHECK(row[r + 1] >= row[r], HFH_ERR_INVALID, "hfh_batch_submit_csr: row pointer decreases at %d", r);
        for (int j = row[r]; j < row[r + 1]; ++j)
            HFH_CHECK(col[j] >= 0 && col[j] < n && (j == row[r] || col[j] > col[j - 1]), HFH_ERR_INVALID,
                      "hfh_batch_submit_csr: row %d: columns must be ascending and in [0, %d)", r, n);
    }
    std::vector<int> sptr;
    const long long E = sell_entries(n, row, &sptr);
    HFH_CHECK(E < (1LL << 31), HFH_ERR_UNSUPPORTED, "hfh_batch_submit_csr: SELL copy exceeds 2^31 entries");
    const int nsl = (n + 63) / 64;
    const Layout L(n, nsl, E, tiles_of(nsl));
    // diagonal of every row (the inner Jacobi preconditioner needs it)
    std::vector<float> invd(n);
    for (int r = 0; r < n; ++r) {
        float d = 0.0f;
        for (int j = row[r]; j < row[r + 1]; ++j)
            if (col[j] == r) d = static_cast<float>(val[j]);
        HFH_CHECK(d != 0.0f, HFH_ERR_ZERO_PIVOT, "hfh_batch_submit_csr: row %d has no nonzero diagonal", r);
        invd[r] = 1.0f / d;
    }
    int slot;
    size_t off;
    int rc;
    if ((rc = reserve(B, L.total, &slot, &off))) return rc;
    auto undo = [&](int code) {
        B->arena.release(off, L.total);
        return code;
    };
    if (hipEventSynchronize(B->ev_up) != hipSuccess) return undo(hip_fail(hipErrorUnknown, "hfh_batch_submit_csr"));
    if ((rc = ensure_stage(B, L.upload))) return undo(rc);
    char *h = static_cast<char *>(B->stage);
    std::memcpy(h + L.ptr, sptr.data(), sizeof(int) * sptr.size());
    int *hl = reinterpret_cast<int *>(h + L.len);
    int *hc = reinterpret_cast<int *>(h + L.col);
    double *h64 = reinterpret_cast<double *>(h + L.v64);
    float *h32 = reinterpret_cast<float *>(h + L.v32);
    for (int s = 0; s < nsl; ++s) {
        const int base = sptr[s], W = (sptr[s + 1] - base) >> 6;
        for (int lane = 0; lane < 64; ++lane) {
            const int r = s * 64 + lane;
            const int j0 = r < n ? row[r] : 0, len = r < n ? row[r + 1] - row[r] : 0;
            if (r < n) hl[r] = len;
            for (int e = 0; e < W; ++e) {
                const int idx = base + e * 64 + lane;
                hc[idx] = e < len ? col[j0 + e] : std::min(r, n - 1);
                h64[idx] = e < len ? val[j0 + e] : 0.0;
                h32[idx] = e < len ? static_cast<float>(val[j0 + e]) : 0.0f;
            }
        }
    }
    std::memcpy(h + L.invd, invd.data(), sizeof(float) * n);
    std::memcpy(h + L.b, b, sizeof(double) * n);
    if (x0) std::memcpy(h + L.x, x0, sizeof(double) * n);
    else std::memset(h + L.x, 0, sizeof(double) * n);
    if (hipMemcpyAsync(B->arena_d + off, h, L.upload, hipMemcpyHostToDevice, B->cs) != hipSuccess ||
        hipEventRecord(B->ev_up, B->cs) != hipSuccess) {
        set_error("hfh_batch_submit_csr: upload failed");
        return undo(HFH_ERR_HIP);
    }
    if ((rc = activate(B, slot, off, L, n, nsl, L.total, 0, id))) return undo(rc);
    return HFH_OK;
}

extern "C" int hfh_batch_submit_fv(hfh_batch *B, int nx, int ny, int nz, const double *kcell, const double *b,
                                   const double *x0, int *id) {
    HFH_CHECK(B && kcell && b && id, HFH_ERR_INVALID, "hfh_batch_submit_fv: bad argument");
    int n, W, rc;
    if ((rc = fv_shape(nx, ny, nz, &n, &W))) return rc;
    for (int c = 0; c < n; ++c)
        HFH_CHECK(kcell[c] > 0 && std::isfinite(kcell[c]), HFH_ERR_INVALID,
                  "hfh_batch_submit_fv: cell coefficient %d must be positive and finite", c);
    const int nsl = (n + 63) / 64;
    const long long E = 64LL * W * nsl;
    HFH_CHECK(E < (1LL << 31), HFH_ERR_UNSUPPORTED, "hfh_batch_submit_fv: SELL copy exceeds 2^31 entries");
    const Layout L(n, nsl, E, tiles_of(nsl), fv_words(W, nsl));
    int slot;
    size_t off;
    if ((rc = reserve(B, L.total, &slot, &off))) return rc;
    auto undo = [&](int code) {
        B->arena.release(off, L.total);
        return code;
    };
    // staging: k (kept in the r64 section until the assembly has run), b, x
    const size_t kbytes = sizeof(double) * n;
    if (hipEventSynchronize(B->ev_up) != hipSuccess) return undo(hip_fail(hipErrorUnknown, "hfh_batch_submit_fv"));
    if ((rc = ensure_stage(B, 3 * up(kbytes)))) return undo(rc);
    char *h = static_cast<char *>(B->stage);
    std::memcpy(h, kcell, kbytes);
    std::memcpy(h + up(kbytes), b, kbytes);
    if (x0) std::memcpy(h + 2 * up(kbytes), x0, kbytes);
    else std::memset(h + 2 * up(kbytes), 0, kbytes);
    hipStream_t cs = B->cs;
    char *base = B->arena_d + off;
    if (hipMemcpyAsync(base + L.r64, h, kbytes, hipMemcpyHostToDevice, cs) != hipSuccess ||
        hipMemcpyAsync(base + L.b, h + up(kbytes), kbytes, hipMemcpyHostToDevice, cs) != hipSuccess ||
        hipMemcpyAsync(base + L.x, h + 2 * up(kbytes), kbytes, hipMemcpyHostToDevice, cs) != hipSuccess) {
        set_error("hfh_batch_submit_fv: upload failed");
        return undo(HFH_ERR_HIP);
    }
    const int threads = std::max(nsl * 64, nsl + 1);
    hipLaunchKernelGGL(kb_fv_assemble, dim3((threads + kBlock - 1) / kBlock), dim3(kBlock), 0, cs, nx, ny, nz, n, nsl,
                       W, reinterpret_cast<const double *>(base + L.r64), reinterpret_cast<int *>(base + L.ptr),
                       reinterpret_cast<int *>(base + L.len), reinterpret_cast<unsigned *>(base + L.col),
                       reinterpret_cast<int *>(base + L.wptr), reinterpret_cast<int *>(base + L.dict),
                       reinterpret_cast<double *>(base + L.v64), reinterpret_cast<float *>(base + L.v32),
                       reinterpret_cast<float *>(base + L.invd));
    if ((rc = launch_check("kb_fv_assemble"))) return undo(rc);
    if (hipEventRecord(B->ev_up, cs) != hipSuccess) return undo(hip_fail(hipErrorUnknown, "hfh_batch_submit_fv"));
    if ((rc = activate(B, slot, off, L, n, nsl, L.total, nz > 1 ? 7 : 5, id))) return undo(rc);
    return HFH_OK;
}

extern "C" int hfh_batch_step(hfh_batch *B, int *active, int *nfinished) {
    HFH_CHECK(B, HFH_ERR_INVALID, "hfh_batch_step: null engine");
    hipStream_t s = B->ctx->stream;
    int rc;
    if (B->tiles_dirty && (rc = upload_tiles(B))) return rc;
    int nact = 0;
    for (const auto &S : B->slots)
        if (S.id >= 0 && !S.reported) ++nact;
    int launched = -1;
    if (nact) {
        launched = B->rb_next;
        B->rb_next ^= 1;
        HFH_HIP(hipEventRecord(B->ev_g0[launched], s));
        // the small systems' `chunk` iterations on the second stream, after
        // the table uploads and slot writes on this one; joined below, so
        // every later slot write on this stream follows them
        if (B->nsmall_now > 0) {
            const KParams kp{B->prm.rtol, B->prm.atol, B->prm.inner_rtol, B->prm.outer_maxits,
                             B->prm.inner_maxits};
            HFH_HIP(hipEventRecord(B->ev_fork, s));
            HFH_HIP(hipStreamWaitEvent(B->ss, B->ev_fork, 0));
            hipLaunchKernelGGL(kb_small, dim3(B->nsmall_now), dim3(kSmallBlock), 0, B->ss, B->d_sys, B->d_st,
                               B->d_small, B->nsmall_now, B->prm.chunk * kSmallChunks, kp);
            if ((rc = launch_check("kb_small"))) return rc;
            HFH_HIP(hipEventRecord(B->ev_join, B->ss));
        }
        if (B->ntiles_now > 0) {
            if (B->exec) {
                HFH_HIP(hipGraphLaunch(B->exec, s));
            } else if ((rc = enqueue_iterations(B, B->prm.chunk))) {
                return rc;
            }
        }
        if (B->nsmall_now > 0) HFH_HIP(hipStreamWaitEvent(s, B->ev_join, 0));
        HFH_HIP(hipEventRecord(B->ev_g1[launched], s));
        ++B->steps;
        HFH_HIP(hipMemcpyAsync(B->h_st[launched], B->d_st, sizeof(SysState) * B->slots.size(),
                               hipMemcpyDeviceToHost, s));
        HFH_HIP(hipEventRecord(B->ev_rb[launched], s));
    }
    // read the previous replay's states while this one runs
    int fin = 0;
    if (B->rb_pending >= 0) {
        fin = take_readback(B, B->rb_pending);
        if (fin < 0) return fin;
        B->rb_pending = -1;
    }
    if (launched >= 0) B->rb_pending = launched;
    if (active) {
        int a = 0;
        for (const auto &S : B->slots)
            if (S.id >= 0 && !S.reported) ++a;
        *active = a;
    }
    if (nfinished) *nfinished = fin;
    return HFH_OK;
}

extern "C" int hfh_batch_finished(hfh_batch *B, int *ids, int cap, int *count) {
    HFH_CHECK(B && count && (ids || cap == 0) && cap >= 0, HFH_ERR_INVALID, "hfh_batch_finished: bad argument");
    const int m = std::min<int>(cap, static_cast<int>(B->finished.size()));
    for (int i = 0; i < m; ++i) ids[i] = B->finished[i];
    *count = static_cast<int>(B->finished.size());
    return HFH_OK;
}

extern "C" int hfh_batch_collect(hfh_batch *B, int id, double *x, hfh_batch_result *res) {
    HFH_CHECK(B, HFH_ERR_INVALID, "hfh_batch_collect: null engine");
    auto it = B->slot_of.find(id);
    HFH_CHECK(it != B->slot_of.end(), HFH_ERR_INVALID, "hfh_batch_collect: unknown request id %d", id);
    const int slot = it->second;
    hfh_batch::Slot &S = B->slots[slot];
    HFH_CHECK(S.reported, HFH_ERR_INVALID, "hfh_batch_collect: request %d has not finished", id);
    // a finished system is never written again, so its solution can be read
    // on the copy stream while later replays run
    if (x) {
        HFH_HIP(hipMemcpyAsync(x, S.x, sizeof(double) * S.n, hipMemcpyDeviceToHost, B->cs));
        HFH_HIP(hipStreamSynchronize(B->cs));
    }
    const SysState &st = S.final_state;
    if (res) {
        res->outer_iters = st.outer;
        res->inner_iters = st.inner_total;
        res->state = st.status;
        res->res = st.res;
        res->res0 = st.res0;
    }
    const int status = st.status;
    // back to P_IDLE behind the running replays; until then the slot stays
    // P_DONE, which every kernel skips, so its region can be reused at once
    hipLaunchKernelGGL(kb_set_slot, dim3(1), dim3(64), 0, B->ctx->stream, B->d_sys, B->d_st, slot, 0, SysDesc{},
                       SysState{});
    const int rc = launch_check("kb_set_slot");
    B->arena.release(S.off, S.bytes);
    B->slot_of.erase(it);
    B->finished.erase(std::remove(B->finished.begin(), B->finished.end(), id), B->finished.end());
    S = hfh_batch::Slot();
    B->tiles_dirty = true;
    if (rc) return rc;
    return status == HFH_SUCCESS_ABS || status == HFH_SUCCESS_REL ? HFH_OK : HFH_ERR_NOT_CONVERGED;
}

extern "C" int hfh_batch_stats(const hfh_batch *B, long long *replays, double *busy_ms, long long *row_iters) {
    HFH_CHECK(B, HFH_ERR_INVALID, "hfh_batch_stats: null engine");
    if (replays) *replays = B->replays;
    if (busy_ms) *busy_ms = B->busy_ms;
    if (row_iters) *row_iters = B->row_iters;
    return HFH_OK;
}

extern "C" int hfh_batch_usage(const hfh_batch *B, int *active, long long *used, long long *cap) {
    HFH_CHECK(B, HFH_ERR_INVALID, "hfh_batch_usage: null engine");
    int a = 0;
    for (const auto &S : B->slots)
        if (S.id >= 0) ++a;
    if (active) *active = a;
    if (used) *used = static_cast<long long>(B->arena.used);
    if (cap) *cap = static_cast<long long>(B->arena.cap);
    return HFH_OK;
}
