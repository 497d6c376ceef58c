// hfh_internal.h -- private definitions of the gfx950 HIP platform library.
//
// Layout in HBM (all arrays 256-byte aligned by hipMalloc):
//   vector  : T[n] contiguous
//   CSR     : int32 row[nrow+1], int32 col[nnz] (ascending per row), T val[nnz],
//             int32 rblk[nrblk+1] row-block boundaries for the streaming SpMV,
//             int32 diag[nrow] position of the diagonal entry (-1 if absent)
//   reductions: per-context scratch of FP64 workgroup partials
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/hfh.h"

namespace hfh {

void set_error(const char *fmt, ...);

// Kernel launch geometry shared by the library.
constexpr int kBlock = 256;          // threads per workgroup (4 wave64)
constexpr int kSpmvTile = 1024;      // nnz staged per workgroup by the streaming SpMV
constexpr int kRedBlocks = 1024;     // max partials per reduction
// Dictionary-coded SELL copies (SpMV, MC-SGS sweeps): the code of a padding
// slot (a slot past its row's length), so kernels need no row lengths;
// dictionaries therefore hold at most 255 offsets
constexpr unsigned kPadCode = 255;
constexpr int kNumXcd = 8;           // MI355X: 8 accelerator complex dies
// Size of the per-device static array that the allocation-free CG
// (hfh_pcg_solve_noalloc) works in: FP64 systems up to 2,096,000 rows.
constexpr size_t kStaticScratch = static_cast<size_t>(64) << 20;

struct Timing {
    bool on = false;
    long long launches[4] = {0, 0, 0, 0};
    double ms[4] = {0, 0, 0, 0};
};

}  // namespace hfh

struct hfh_ctx {
    int device = 0;
    int num_cu = 0;
    long long hbm = 0;
    char name[128] = {0};
    hipStream_t stream = nullptr;
    // reduction scratch: partials[kRedBlocks * 8] doubles, then result slots
    double *red = nullptr;            // device
    double *host_res = nullptr;       // pinned host mirror for results
    hfh::Timing timing;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int live = 0;                     // vectors, matrices and preconditioners created on the context
    // device address of this device's static scratch array (hfh_krylov.hip)
    void *scratch = nullptr;
};

struct hfh_vec {
    hfh_ctx *ctx = nullptr;
    int n = 0;
    int dtype = HFH_F64;
    void *d = nullptr;
    // index set (lVector::set_indexset): device copy for the gather / scatter
    // of GetIndexedValues / SetIndexedValues, host copy for get_indexset
    int *iset = nullptr;
    std::vector<int> h_iset;
    size_t elem() const { return dtype == HFH_F64 ? 8 : 4; }
};

struct hfh_csr {
    hfh_ctx *ctx = nullptr;
    int nrow = 0, ncol = 0, nnz = 0, dtype = HFH_F64;
    int *row = nullptr, *col = nullptr;
    void *val = nullptr;
    int *rblk = nullptr;   // streaming row blocks
    int nrblk = 0;
    int spmv_algo = HFH_SPMV_STREAM, spmv_p1 = 2048;   // configured kernel
    int spmv_algo_req = 0, spmv_p1_req = 0, spmv_p2_req = 0;  // requested (0 = automatic)
    int nwg = 1;           // SpMV workgroups = partial sums of a fused dot
    std::vector<int> h_col;    // host copy of the columns, kept for the analyses
    // SELL-64 copy for HFH_SPMV_SELL: slices of 64 rows stored column-major
    int nslice = 0;
    int *sl_ptr = nullptr;     // [nslice+1] element offset of each slice
    int *sl_len = nullptr;     // [nrow] row lengths
    int *sl_col = nullptr;     // padded columns (padding points at the row itself)
    void *sl_val = nullptr;    // padded values (padding 0)
    bool sell_dirty = false;   // SELL copy must be rebuilt from the CSR arrays
    // HFH_SPMV_SELL_DICT: columns as one-byte codes into a table of distinct
    // column offsets (col - row); four codes of consecutive slots per word
    int *sd_dict = nullptr;    // [sd_ndict] offsets
    int sd_ndict = 0;
    unsigned *sd_code = nullptr;   // code words, column-major per slice
    int *sd_wptr = nullptr;    // [nslice+1] word offset of each slice
    int sd_code0 = 0;          // code of offset 0 (the diagonal)
    // HFH_SPMV_SELL_LDS: the pair layout's values with each slot's column as a
    // 16-bit index into its slice's LDS image of x (the slice's distinct
    // chunks of x, ascending): chunk position * sx_chunk + column % sx_chunk
    unsigned short *sx_code = nullptr;   // [slots] pair layout, padding 0
    unsigned short *sx_dcode = nullptr;  // [nrow] code of the diagonal (0xffff: none)
    int *sx_cptr = nullptr;    // [nslice+1] offset of each slice's chunk list
    int *sx_cid = nullptr;     // chunk ids (x index / sx_chunk)
    int sx_chunk = 0;          // elements per chunk (16)
    int sx_img = 0;            // largest image of a slice (chunks * sx_chunk)
    int *diag = nullptr;   // diagonal position per row
    bool has_structure = false;
    bool sorted = true;      // columns ascending in every row (false after a diagonal-first swap)
    unsigned version = 0;    // bumped whenever the values or the structure change
    std::vector<int> h_row;  // host copy of the row pointer (for analysis)
    size_t elem() const { return dtype == HFH_F64 ? 8 : 4; }
};

// Level schedule on the device: rows in level order, level boundaries on host.
struct hfh_levels {
    int nlev = 0;
    std::vector<int> ptr;      // host, nlev+1
    int *rows = nullptr;       // device, nrow
    int *ptr_d = nullptr;      // device copy of ptr
    std::vector<std::pair<int, int>> runs;   // launch plan: level ranges [first, last)
};

// Rows of a preconditioner sweep in SELL-64 form, slices of 64 consecutive
// rows of one colour, column-major, padded to the longest row of the slice.
struct hfh_sweep {
    int nslice = 0;
    int *ptr = nullptr;        // [nslice+1] element offset of each slice
    int *len = nullptr;        // [nrow] entries of each row in this sweep
    int *col = nullptr;        // padded columns (padding: a row of the same slice)
    void *val = nullptr;       // padded values (padding 0)
    std::vector<int> slice0;   // host: first slice of each colour (ncolors+1)
    // with at most 256 distinct offsets col - row over the sweep's entries,
    // the columns are one-byte codes into `dict` instead (col is null):
    // code word w of a lane holds slots 4w..4w+3; the values keep the
    // slice's width, its code words cover it rounded up to 4 slots
    unsigned *code = nullptr;  // code words
    int *wptr = nullptr;       // [nslice+1] word offset of each slice
    int *dict = nullptr;       // [ndict] offsets, ascending (0 included: padding)
    int ndict = 0;
};

// Rows of a level-scheduled triangular sweep, sliced in level order: the rows
// of level l fill slices lslice[l] .. lslice[l+1] (64 rows each, lane = row),
// each slice column-major and padded to its longest row.  A row's entries are
// its lower (forward sweep) or upper (backward sweep) ones in stored order;
// its diagonal value sits beside them.  Values are refilled from the CSR
// arrays through the position maps when the matrix's values change.
struct hfh_tsweep {
    int nslice = 0;
    long long nent = 0;
    std::vector<int> lslice;   // host: first slice of each level (nlev + 1)
    int *lslice_d = nullptr;   // device copy
    int *ptr = nullptr;        // [nslice + 1] element offset of each slice
    int *row = nullptr;        // [64 nslice] row of each lane (-1: padding lane)
    int *len = nullptr;        // [64 nslice] entries of each lane
    int *col = nullptr;        // [nent] columns (padding: 0)
    int *pos = nullptr;        // [nent] CSR position of each entry (-1: padding)
    int *dpos = nullptr;       // [64 nslice] CSR position of each lane's diagonal (-1: padding)
    void *val = nullptr;       // [nent] values
    void *dval = nullptr;      // [64 nslice] diagonal values
    // Blocked form only (slices are geometric blocks of a structured grid,
    // in block-level order): a column below 0 names the lane -(col + 1) of
    // the same block, whose result arrives through a cross-lane shuffle
    int *lev = nullptr;        // [64 nslice] level of each lane inside its block (-1: padding)
    int *bnl = nullptr;        // [nslice] levels inside each block
    int kmax = 0;              // most entries of any lane
};

struct hfh_prec {
    int kind = HFH_PREC_NONE;
    hfh_ctx *ctx = nullptr;
    const hfh_csr *A = nullptr;     // operator (SGS, MC-SGS) or LU factors (ILU0)
    int n = 0;
    int dtype = HFH_F64;
    void *invd = nullptr;           // Jacobi / MC-SGS: 1/a_ii
    void *dv = nullptr;             // MC-SGS: -a_ii / (omega (2 - omega))
    void *sval = nullptr;           // MC-SGS: omega-scaled values (nullptr when omega == 1)
    double omega = 1.0;
    int ncolors = 0;
    std::vector<int> color_off;     // host, ncolors+1
    hfh_levels lo, up;              // SGS / ILU0 level schedules
    hfh_sweep fw, bw;               // MC-SGS: lower-colour / upper-colour entries per row
    hfh_tsweep tlo, tup;            // SGS / ILU0: forward / backward sweep rows by level
    hfh_tsweep blo, bup;            // SGS / ILU0 on a structured grid: the sweeps by geometric block
    unsigned tversion = 0;          // A->version the sweep values were copied at
    // SGS / ILU0 outside the fused drivers: both sweeps captured once as a
    // hipGraph on internal in/out buffers (one launch instead of one per level)
    void *gin = nullptr, *gout = nullptr;
    hipGraph_t sweep_graph = nullptr;
    hipGraphExec_t sweep_exec = nullptr;
    bool graph_failed = false;
    // SGS / ILU0 sweep schedule: sync-free (one launch per sweep) or one
    // launch per level; the sync-free form keeps the forward results in sf_y
    // and its ticket counters and failure word in sf_ctl
    int sweep_mode = HFH_SWEEP_SYNCFREE;
    int sf_waves = 0;                  // sync-free: waves per CU (0: from the level widths)
    void *sf_y = nullptr;
    int *sf_ctl = nullptr;
};

namespace hfh {

inline int hip_fail(hipError_t e, const char *what) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return HFH_ERR_HIP;
}

#define HFH_HIP(call)                                          \
    do {                                                       \
        hipError_t e_ = (call);                                \
        if (e_ != hipSuccess) return hfh::hip_fail(e_, #call); \
    } while (0)

#define HFH_CHECK(cond, code, ...)           \
    do {                                     \
        if (!(cond)) {                       \
            hfh::set_error(__VA_ARGS__);     \
            return (code);                   \
        }                                    \
    } while (0)

int dev_alloc(void **p, size_t bytes);
int launch_check(const char *what);
// False when the environment sets HFH_GRAPHS=0: the drivers and the serving
// engine then launch every iteration directly instead of replaying a
// captured graph (the same kernels in the same order, the same results).
bool graphs_enabled();
// Device address of the static scratch array on `device` (current), looked
// up once per device (hfh_krylov.hip).
int static_scratch(int device, void **ptr);
// HFH_ERR_HIP (and the failure word cleared) when a sync-free sweep of P gave
// up waiting since the last check; HFH_OK otherwise (also for P = NULL).
int sweep_failed(hfh_prec *P);

// Reductions (hfh_blas1.hip).  Results land in ctx->red + slot (device).
int dot_async(hfh_ctx *ctx, int dtype, int n, const void *x, const void *y, double *dst);
int fetch(hfh_ctx *ctx, const double *dsrc, double *host, int count);
// Sums the first nb single-quantity partials of ctx->red into dst (device) in
// fixed order (one finisher workgroup).
int partials_sum(hfh_ctx *ctx, int nb, double *dst);

// Timing helpers.
struct ScopedKernelTimer {
    hfh_ctx *ctx;
    int kclass;
    bool on;
    ScopedKernelTimer(hfh_ctx *c, int k);
    ~ScopedKernelTimer();
};

// SpMV launch used by the drivers (hfh_spmv.hip).  mode 0: y = A x;
// mode 1: y += A x.  When w is non-null each workgroup also stores its
// partial sum of y_i * w_i at part[row block] (A->nrblk partials), to be
// summed by a finisher kernel (A->nwg partials).
int spmv_launch(const hfh_csr *A, const void *x, void *y, int mode);
int spmv_configure(hfh_csr *A, int algo, int p1, int p2);
int spmv_sell_fill(hfh_csr *A);
// After the device columns changed order: diagonal positions, SpMV layout
// and SELL copy rebuilt from the CSR arrays (host copies already updated).
int spmv_refresh_structure(hfh_csr *A);
int spmv_launch_ex(const hfh_csr *A, const void *x, void *y, int mode, const void *w, double *part,
                   const int *done);
// One Jacobi step xn = update(b - sum_{j != r} a_rj x_j) over the LDS-staged
// or the dictionary-coded SELL copy (A->spmv_algo HFH_SPMV_SELL_LDS or
// HFH_SPMV_SELL_DICT); T is A's value type.
template <typename T>
int spmv_coded_jacobi(const hfh_csr *A, const T *x, T *xn, const T *b, const T *invd, bool damped, T w, T omw);

// Preconditioner apply on raw device pointers (hfh_prec.hip).  `done`
// (device, nullable): every kernel returns at once when *done != 0.
// mc_first_done: for a multicolour SGS of two or more colours, the first
// forward step (colour 0) is already in `out` (the CG update computes it).
// rzp / rz_out (multicolour SGS only): the apply also forms in.out (r.z of a
// CG step) through partials in rzp (mcsgs_rz_slots(P) doubles) into *rz_out.
int prec_apply_raw(hfh_prec *P, const void *in, void *out, const int *done = nullptr, bool mc_first_done = false,
                   double *rzp = nullptr, double *rz_out = nullptr);
int mcsgs_rz_slots(const hfh_prec *P);
// Brings a preconditioner's copies of the matrix values up to date (after the
// matrix's values changed); a no-op while the stream is being captured.
int prec_sync(hfh_prec *P);

// XCD-aware workgroup remap: consecutive logical tiles land on one XCD so its
// L2 keeps the shared gather window.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int per = nb / kNumXcd, rem = nb % kNumXcd;
    const int xcd = b % kNumXcd, idx = b / kNumXcd;
    // XCDs [0, rem) own per+1 tiles, the rest own per tiles
    const int base = xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
    return base + idx;
}

// 16-byte vectors of a stream: 2 doubles or 4 floats.
template <typename T> struct V16;
template <> struct V16<double> {
    typedef double type __attribute__((ext_vector_type(2)));
    static constexpr int W = 2;
};
template <> struct V16<float> {
    typedef float type __attribute__((ext_vector_type(4)));
    static constexpr int W = 4;
};

// Separately rounded product and sum.  The contraction pragma sits inside
// the helpers so that it also holds where they are inlined into code
// compiled with contraction on.
template <typename T>
__device__ __forceinline__ T mul_rn(T a, T b) {
#pragma clang fp contract(off)
    return a * b;
}
template <typename T>
__device__ __forceinline__ T add_rn(T a, T b) {
#pragma clang fp contract(off)
    return a + b;
}

// Jacobi step epilogue for row r given s = sum_{j != r} a_rj x_j
// (DoNormalIteration / DoDampedIteration, linear_solver/jacobi.cc):
// y = (b_r - s) * inv_diag_r; damped: (1 - w) x_r + w y, each step rounded.
template <typename T, bool DAMPED>
__device__ __forceinline__ T jacobi_update(T s, T xr, T br, T dr, T w, T omw) {
    const T y = mul_rn(add_rn(br, -s), dr);
    if (!DAMPED) return y;
    return add_rn(mul_rn(xr, omw), mul_rn(w, y));
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

// Deterministic block sum of kBlock threads; result valid in thread 0.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *sh /* >= kBlock/64 */) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    T s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / 64; ++i) s += sh[i];
    return s;
}

// Cross-workgroup reductions never synchronise inside a kernel: on gfx950
// an agent-scope release fence writes back the XCD's L2, so a per-workgroup
// fence + atomic costs far more than the reduction itself.  Each workgroup
// stores its partial sums with plain stores; a finisher kernel launched
// next on the same stream (the kernel boundary makes the partials visible)
// sums them in a fixed order, so results are bit-reproducible.
template <int NQ>
__device__ __forceinline__ void store_partials(const double (&v)[NQ], double *part, int stride, int slot) {
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) part[q * stride + slot] = v[q];
    }
}

constexpr int kFinBlock = 1024;  // threads of a finisher workgroup

// Sums part[q*stride + b], b < nb, with one workgroup of kFinBlock threads:
// lane-strided sums in ascending b, then a fixed tree.  Valid in thread 0.
template <int NQ>
__device__ __forceinline__ void reduce_partials(const double *part, int stride, int nb, double (&out)[NQ]) {
    __shared__ double sh[kFinBlock / 64];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double s = 0;
        // unrolled so that the loads issue together; the adds keep their order
#pragma unroll 8
        for (int b = threadIdx.x; b < nb; b += kFinBlock) s += part[q * stride + b];
        s = wave_sum(s);
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        if (lane == 0) sh[w] = s;
        __syncthreads();
        double t = 0;
        if (threadIdx.x == 0)
            for (int i = 0; i < kFinBlock / 64; ++i) t += sh[i];
        out[q] = t;
        __syncthreads();
    }
}

// The sum reduce_partials<1> returns, computed by a kBlock-thread workgroup
// (every workgroup of a kernel that needs it, in place of a finisher
// launch).  Thread t stands in for the finisher's threads t + kBlock k,
// k < kFinBlock / kBlock: those sit in lane t mod 64 of the finisher's waves
// t / 64 + (kBlock / 64) k, so each stand-in's strided sum and wave tree are
// the finisher's own and the wave totals meet in the same order: the result
// has the finisher's bits.  Valid in every thread.
__device__ __forceinline__ double fin_sum_all(const double *__restrict__ part, int nb,
                                              double *sh /* kFinBlock/64 */, double *bc) {
    constexpr int K = kFinBlock / kBlock;
    double s[K];
#pragma unroll
    for (int k = 0; k < K; ++k) s[k] = 0;
    for (int b0 = 0; b0 < nb; b0 += kFinBlock) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int b = b0 + static_cast<int>(threadIdx.x) + kBlock * k;
            if (b < nb) s[k] += part[b];
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double t = wave_sum(s[k]);
        if (lane == 0) sh[w + (kBlock / 64) * k] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int i = 0; i < kFinBlock / 64; ++i) t += sh[i];
        *bc = t;
    }
    __syncthreads();
    const double v = *bc;
    __syncthreads();
    return v;
}

}  // namespace hfh
