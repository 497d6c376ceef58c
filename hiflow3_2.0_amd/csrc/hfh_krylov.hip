// hfh_krylov.hip -- device-resident Krylov drivers.
//
// CG (linear_solver/cg.cc:163-248) is restated with its scalars kept in HBM so
// the host never waits on a Dot.  One Jacobi iteration is three or four launches:
//   K1  Ap = A p with per-workgroup partial sums of Ap.p (summed by a
//       finisher when the folding below would re-read too many partials)
//   K2  every workgroup sums the pAp partials (the same fixed order in each);
//       alpha = beta/pAp; r -= alpha Ap; partials of r.r (and r.z with
//       z = r.*invD for Jacobi); workgroup 0 advances the step
//   K3  x += alpha p (one read of p serves this and the direction update);
//       every workgroup sums the r.r / r.z partials and applies
//       IterateControl::Check to res = sqrt(r.r); workgroup 0 records the
//       step and the new beta; unless finished, p = beta_ratio p + z
// With another preconditioner, [M] z = M^-1 r and K2b (partials of r.z) run
// between K2 and K3; for a multicolour SGS, K2 also writes the colour-0
// forward step of [M], and the sweeps that finalise z store the r.z partials
// in place of K2b (a one-workgroup finisher sums them).  No kernel reads a state field that another
// workgroup of the same launch writes (beta_prev keeps K2's beta for K3).
// Every kernel returns at once when the solve is done (K3 on the idle flag
// that K2 sets then), so the host enqueues chunks of iterations (a captured
// hipGraph when possible) and reads the state once per chunk.  Compulsory
// traffic per Jacobi iteration: 12 nnz + 100 n bytes (FP64).
#include "hfh_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>

namespace hfh {

struct KState {
    double pap, rr, rz;     // reduction results
    double beta, ratio;     // beta = r.z (or r.r), ratio = new/old
    double beta_prev;       // beta of the step in progress (read by K3 while workgroup 0 updates beta)
    double res, res0;
    double atol, rtol, dtol;
    int iter, done, state, maxits;
    int idle;               // set by K2 when it finds the solve done; K3 returns on it
};

__device__ __forceinline__ int check_state(int step, double v, double first, const KState *s) {
    if (v < s->atol) return HFH_SUCCESS_ABS;
    if (v < s->rtol * first) return HFH_SUCCESS_REL;
    if (v > s->dtol * first) return HFH_FAIL_DIV;
    if (step >= s->maxits) return HFH_FAIL_MAXITS;
    return HFH_ITERATE;
}

// Sum of part[0..nb) by one workgroup in a fixed order (lane-strided sums,
// then the block tree), returned to every thread.  Every workgroup that calls
// it on the same partials gets the same value.
__device__ __forceinline__ double block_sum_all(const double *__restrict__ part, int nb, double *sh, double *bc) {
    double v = 0;
    // unrolled so that the loads issue together; the adds keep their order
#pragma unroll 4
    for (int b = threadIdx.x; b < nb; b += kBlock) v += part[b];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) *bc = v;
    __syncthreads();
    v = *bc;
    __syncthreads();
    return v;
}

// The sums of pa[0..na) and pb[0..nb) that two block_sum_all calls return
// (the same order of adds for each), with the loads of both in one pass and
// one set of barriers.
__device__ __forceinline__ void block_sum_all2(const double *__restrict__ pa, int na, const double *__restrict__ pb,
                                               int nb, double *sh /* 2 kBlock/64 */, double *bc /* 2 */, double &a,
                                               double &b) {
    double va = 0, vb = 0;
    const int m = max(na, nb);
#pragma unroll 4
    for (int i = threadIdx.x; i < m; i += kBlock) {
        if (i < na) va += pa[i];
        if (i < nb) vb += pb[i];
    }
    va = wave_sum(va);
    vb = wave_sum(vb);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sh[w] = va;
        sh[kBlock / 64 + w] = vb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ta = 0, tb = 0;
        for (int k = 0; k < kBlock / 64; ++k) ta += sh[k];
        for (int k = 0; k < kBlock / 64; ++k) tb += sh[kBlock / 64 + k];
        bc[0] = ta;
        bc[1] = tb;
    }
    __syncthreads();
    a = bc[0];
    b = bc[1];
    __syncthreads();
}

static inline int host_check(int step, double v, double first, const hfh_control &c) {
    if (v < c.atol) return HFH_SUCCESS_ABS;
    if (v < c.rtol * first) return HFH_SUCCESS_REL;
    if (v > c.dtol * first) return HFH_FAIL_DIV;
    if (step >= c.maxits) return HFH_FAIL_MAXITS;
    return HFH_ITERATE;
}

// MODE 0: no preconditioner (z = r), 1: Jacobi (z = r.*invd), 2: general (z
// computed by a separate apply; K2 reduces r.r only).
// K2: r -= alpha Ap, partials of r.r and (MODE 1) r.(invD r).  The step's
// x += alpha p runs in K3, which reads p anyway (one pass over p fewer).
// The vectors are whole device allocations: nv 16-byte vectors of each are
// read as such, the n mod W elements after them one by one.
template <typename T, int MODE>
__device__ __forceinline__ void pcg_update_one(T &r, T a, T d, T alpha, double &rr, double &rz) {
    const T ri = r + (-alpha) * a;
    r = ri;
    rr += static_cast<double>(ri) * static_cast<double>(ri);
    if (MODE == 1) {
        const T zi = ri * d;
        rz += static_cast<double>(ri) * static_cast<double>(zi);
    }
}

// F0 (MODE 2 with a multicolour SGS preconditioner): also the
// preconditioner's first forward step, z_i = (-r_i) * idv_i for the rows
// i < n0 of colour 0 (k_mc_fwd0's arithmetic), from the new r in registers.
template <typename T, int MODE, bool F0>
__global__ __launch_bounds__(kBlock) void k_pcg_update(int n, long long nv, T *__restrict__ r, const T *__restrict__ Ap,
                                                       const T *__restrict__ invd, double *part, int stride,
                                                       KState *st, T *__restrict__ z0,
                                                       const T *__restrict__ idv0, int n0,
                                                       const double *__restrict__ spart, int nsp) {
    // K3 cannot test `done` itself: its workgroup 0 sets it on the step that
    // ends the solve while later workgroups of that launch still run the
    // step's x update.  K2 sees `done` only on the launches after that step.
    if (st->done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) st->idle = 1;
        return;
    }
    typedef typename V16<T>::type V;
    constexpr int W = V16<T>::W;
    __shared__ double sh[kBlock / 64];
    __shared__ double bc;
    // pAp: with nsp > 0 every workgroup sums the SpMV's nsp partials in the
    // same fixed order (no finisher launch), else k_fin_pap has stored it;
    // workgroup 0 advances the step count and keeps beta for K3, which reads
    // beta_prev while it updates beta
    const double pap = nsp > 0 ? block_sum_all(spart, nsp, sh, &bc) : st->pap;
    const double beta_d = st->beta;
    const T alpha = static_cast<T>(static_cast<T>(beta_d) / static_cast<T>(pap));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (nsp > 0) st->pap = pap;
        st->beta_prev = beta_d;
        st->iter = st->iter + 1;
    }
    double rr = 0, rz = 0;
    const long long step = static_cast<long long>(gridDim.x) * kBlock;
    const long long t0 = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x;
    V *__restrict__ rv = reinterpret_cast<V *>(r);
    const V *__restrict__ av = reinterpret_cast<const V *>(Ap);
    const V *__restrict__ dv = reinterpret_cast<const V *>(invd);
    for (long long i = t0; i < nv; i += step) {
        V ri = rv[i];
        const V ai = av[i];
        V di{};
        if (MODE == 1) di = dv[i];
#pragma unroll
        for (int w = 0; w < W; ++w) {
            T rw = ri[w];
            pcg_update_one<T, MODE>(rw, ai[w], MODE == 1 ? di[w] : T(0), alpha, rr, rz);
            ri[w] = rw;
        }
        rv[i] = ri;
        if (F0) {
            const long long e0 = i * W;
            if (e0 + W <= n0) {
                const V dd = reinterpret_cast<const V *>(idv0)[i];
                V zz;
#pragma unroll
                for (int w = 0; w < W; ++w) zz[w] = (ri[w] * static_cast<T>(-1)) * dd[w];
                reinterpret_cast<V *>(z0)[i] = zz;
            } else if (e0 < n0) {
#pragma unroll
                for (int w = 0; w < W; ++w)
                    if (e0 + w < n0) z0[e0 + w] = (ri[w] * static_cast<T>(-1)) * idv0[e0 + w];
            }
        }
    }
    for (long long j = nv * W + t0; j < n; j += step) {
        pcg_update_one<T, MODE>(r[j], Ap[j], MODE == 1 ? invd[j] : T(0), alpha, rr, rz);
        if (F0 && j < n0) z0[j] = (r[j] * static_cast<T>(-1)) * idv0[j];
    }
    double v[2];
    v[0] = block_sum(rr, sh);
    __syncthreads();
    v[1] = MODE == 1 ? block_sum(rz, sh) : 0.0;
    store_partials<2>(v, part, stride, blockIdx.x);
}

// stores pAp = sum of the SpMV partials (when K2 does not sum them itself)
__global__ __launch_bounds__(kFinBlock) void k_fin_pap(const double *part, int nb, KState *st) {
    if (st->done) return;
    double out[1];
    reduce_partials<1>(part, 0, nb, out);
    if (threadIdx.x == 0) st->pap = out[0];
}

// K2 sums the SpMV partials itself when that re-reads at most this many
// partials over all its workgroups (cheaper than a finisher launch: the
// 2D Q1 256^2 CG step 17.2 -> 14.0 us); above it, the finisher k_fin_pap
// (the 3D Q1 128^3 SpMV's 8,008 partials times 1,024 workgroups cost more
// than the launch they save).
constexpr long long kFoldReads = 1 << 20;

// partials of r.z after a general preconditioner apply
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_rz(int n, const T *__restrict__ r, const T *__restrict__ z,
                                                   double *part, int stride, const KState *st) {
    if (st->done) return;
    __shared__ double sh[kBlock / 64];
    double rz = 0;
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        rz += static_cast<double>(r[i]) * static_cast<double>(z[i]);
    const double v[1] = {block_sum(rz, sh)};
    store_partials<1>(v, part, stride, blockIdx.x);
}


// p = ratio * p + z  (ScaleAdd, cg.cc:233).  MODE as in k_pcg_update; nv
// as there.
template <typename T, int MODE>
__device__ __forceinline__ T pcg_dir_one(T p, T r, T z, T d, T ratio) {
    T zi;
    if (MODE == 0) zi = r;
    else if (MODE == 1) zi = r * d;
    else zi = z;
    return ratio * p + zi;
}

// K3: x += alpha p (cg.cc:213, with K2's alpha = beta_prev / pAp), then the
// step's bookkeeping (IterateControl::Check on the recursive residual,
// cg.cc:200-231): every workgroup sums the partials of r.r (slot 0) and r.z
// (slot `stride`, MODE != 0) of K2 / K2b in the same fixed order and takes
// the same decision; workgroup 0 records the step (residual, history, state,
// done, the new beta).  A finished solve updates x and leaves p alone.
// nbz: r.z partials in slot `stride` (1 when a finisher has summed them).
template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void k_pcg_p(int n, long long nv, T *__restrict__ x, T *__restrict__ p,
                                                  const T *__restrict__ r, const T *__restrict__ z,
                                                  const T *__restrict__ invd, KState *st,
                                                  const double *__restrict__ part, int stride, int nb, int nbz,
                                                  double *hist) {
    if (st->idle) return;
    typedef typename V16<T>::type V;
    constexpr int W = V16<T>::W;
    __shared__ double sh[2 * (kBlock / 64)];
    __shared__ double bc[2];
    double trr, trz = 0.0;
    if (MODE != 0)
        block_sum_all2(part, nb, part + stride, nbz, sh, bc, trr, trz);
    else
        trr = block_sum_all(part, nb, sh, bc);
    const int it = st->iter;
    const T rrT = static_cast<T>(trr);
    const double res = static_cast<double>(static_cast<T>(sqrt(static_cast<double>(rrT))));
    const int state = check_state(it, res, st->res0, st);
    const T gamma = MODE == 0 ? rrT : static_cast<T>(trz);
    const T ratio = gamma / static_cast<T>(st->beta_prev);
    const T alpha = static_cast<T>(static_cast<T>(st->beta_prev) / static_cast<T>(st->pap));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->rr = trr;
        st->res = res;
        if (hist) hist[it] = res;
        st->state = state;
        if (state != HFH_ITERATE) {
            st->done = 1;
        } else {
            st->ratio = static_cast<double>(ratio);
            st->beta = static_cast<double>(gamma);
        }
    }
    const bool go = state == HFH_ITERATE;
    const long long step = static_cast<long long>(gridDim.x) * kBlock;
    const long long t0 = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x;
    V *__restrict__ xv = reinterpret_cast<V *>(x);
    V *__restrict__ pv = reinterpret_cast<V *>(p);
    const V *__restrict__ rv = reinterpret_cast<const V *>(r);
    const V *__restrict__ zv = reinterpret_cast<const V *>(z);
    const V *__restrict__ dv = reinterpret_cast<const V *>(invd);
    for (long long i = t0; i < nv; i += step) {
        V pi = pv[i], xi = xv[i];
#pragma unroll
        for (int w = 0; w < W; ++w) xi[w] += alpha * pi[w];
        xv[i] = xi;
        if (!go) continue;
        V ri{}, zi{}, di{};
        if (MODE != 2) ri = rv[i];
        if (MODE == 1) di = dv[i];
        if (MODE == 2) zi = zv[i];
#pragma unroll
        for (int w = 0; w < W; ++w) pi[w] = pcg_dir_one<T, MODE>(pi[w], ri[w], zi[w], di[w], ratio);
        pv[i] = pi;
    }
    for (long long j = nv * W + t0; j < n; j += step) {
        const T pj = p[j];
        x[j] += alpha * pj;
        if (go)
            p[j] = pcg_dir_one<T, MODE>(pj, MODE != 2 ? r[j] : T(0), MODE == 2 ? z[j] : T(0),
                                        MODE == 1 ? invd[j] : T(0), ratio);
    }
}

// Writes the initial solver state and hist[0] from the kernel arguments
// (the allocation-free solve has no pinned staging of its own).
__global__ void k_set_state(KState *st, KState v, double *hist, double res0) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    *st = v;
    if (hist) hist[0] = res0;
}

// Workspace of one solve.
struct Work {
    hfh_ctx *ctx = nullptr;
    std::vector<void *> bufs;
    ~Work() {
        for (void *b : bufs)
            if (b) (void)hipFree(b);
    }
    int alloc(void **p, size_t bytes) {
        int rc = dev_alloc(p, bytes);
        if (!rc) bufs.push_back(*p);
        return rc;
    }
};

static int stream_blocks(hfh_ctx *ctx, int n) {
    long long nb = (static_cast<long long>(n) + kBlock * 4 - 1) / (kBlock * 4);
    const long long cap = std::min<long long>(kRedBlocks, std::max(ctx->num_cu, 1) * 4);
    nb = std::max<long long>(1, std::min(nb, cap));
    return static_cast<int>(nb);
}

template <typename T>
struct PcgRun {
    const hfh_csr *A = nullptr;
    hfh_prec *P = nullptr;
    int mode = 0;  // 0 none, 1 jacobi, 2 general
    int n = 0;
    T *x = nullptr, *r = nullptr, *p = nullptr, *z = nullptr, *Ap = nullptr;
    const T *invd = nullptr;
    double *part = nullptr;
    int stride = 0;
    double *spart = nullptr;      // SpMV partials (A->nwg)
    KState *st = nullptr;
    double *hist = nullptr;
    int nb = 1;
    long long nv = 0;             // 16-byte vectors of the element-wise kernels (0: unaligned operand)
    // multicolour SGS with two or more colours: its first forward step (rows
    // of colour 0, z = -r * idv) runs inside k_pcg_update
    bool f0 = false;
    const T *idv0 = nullptr;
    int n0 = 0;
    // multicolour SGS: the sweeps that finalise z store r.z partials here and
    // the apply sums them into slot `stride` (no separate r.z pass)
    double *rzp = nullptr;
    // the instantiated graph of one chunk of iterations, kept for the next
    // solve on this workspace (the same buffers; re-captured when x or the
    // chunk differs); destroyed with the workspace
    hipGraphExec_t exec = nullptr;
    const T *exec_x = nullptr;
    int exec_chunk = 0;

    // SpMV with fused p.Ap partials: summed by every K2 workgroup (fold) or
    // by the finisher k_fin_pap
    bool fold() const { return static_cast<long long>(A->nwg) * nb <= kFoldReads; }
    int spmv_kernel() { return spmv_launch_ex(A, p, Ap, 0, p, spart, &st->done); }
    int fin_pap() {
        if (fold()) return HFH_OK;
        hipLaunchKernelGGL(k_fin_pap, dim3(1), dim3(kFinBlock), 0, A->ctx->stream, (const double *)spart, A->nwg, st);
        return launch_check("k_fin_pap");
    }

    int tail() {
        hipStream_t s = A->ctx->stream;
        const int nsp = fold() ? A->nwg : 0;
        int rc;
        if (mode == 0)
            hipLaunchKernelGGL((k_pcg_update<T, 0, false>), dim3(nb), dim3(kBlock), 0, s, n, nv, r, Ap, invd, part, stride, st, z, idv0, n0, (const double *)spart, nsp);
        else if (mode == 1)
            hipLaunchKernelGGL((k_pcg_update<T, 1, false>), dim3(nb), dim3(kBlock), 0, s, n, nv, r, Ap, invd, part, stride, st, z, idv0, n0, (const double *)spart, nsp);
        else if (f0)
            hipLaunchKernelGGL((k_pcg_update<T, 2, true>), dim3(nb), dim3(kBlock), 0, s, n, nv, r, Ap, invd, part, stride, st, z, idv0, n0, (const double *)spart, nsp);
        else
            hipLaunchKernelGGL((k_pcg_update<T, 2, false>), dim3(nb), dim3(kBlock), 0, s, n, nv, r, Ap, invd, part, stride, st, z, idv0, n0, (const double *)spart, nsp);
        if ((rc = launch_check("k_pcg_update"))) return rc;
        int nbz = nb;
        if (mode == 2) {
            // general preconditioner: z = M^-1 r and the partials of r.z
            // (slot 1, beside r.r in slot 0), both reduced by K3.  On the
            // step that converges, the apply runs once more than needed.
            if (rzp) {
                if ((rc = prec_apply_raw(P, r, z, &st->done, f0, rzp, part + stride))) return rc;
                nbz = 1;
            } else {
                if ((rc = prec_apply_raw(P, r, z, &st->done, f0))) return rc;
                hipLaunchKernelGGL(k_pcg_rz<T>, dim3(nb), dim3(kBlock), 0, s, n, r, z, part + stride, stride, (const KState *)st);
                if ((rc = launch_check("k_pcg_rz"))) return rc;
            }
        }
        const double *pt = part;
        if (mode == 0)
            hipLaunchKernelGGL((k_pcg_p<T, 0>), dim3(nb), dim3(kBlock), 0, s, n, nv, x, p, r, z, invd, st, pt, stride, nb, nbz, hist);
        else if (mode == 1)
            hipLaunchKernelGGL((k_pcg_p<T, 1>), dim3(nb), dim3(kBlock), 0, s, n, nv, x, p, r, z, invd, st, pt, stride, nb, nbz, hist);
        else
            hipLaunchKernelGGL((k_pcg_p<T, 2>), dim3(nb), dim3(kBlock), 0, s, n, nv, x, p, r, z, invd, st, pt, stride, nb, nbz, hist);
        return launch_check("k_pcg_p");
    }

    int iteration() {
        int rc = spmv_kernel();
        if (!rc) rc = fin_pap();
        return rc ? rc : tail();
    }
};

// Runs the chunked iteration loop until the device state reports done.  With
// timing on, each SpMV is bracketed by HIP events on the context stream and
// the launches that did work are added to timing class 0.
template <typename T>
static int pcg_loop(PcgRun<T> &R, int maxits, bool use_graph, bool allow_timing = true) {
    hfh_ctx *c = R.A->ctx;
    hipStream_t s = c->stream;
    const bool timing = c->timing.on && allow_timing;
    use_graph = use_graph && graphs_enabled();
    const int chunk = std::max(1, std::min(16, maxits));
    hipGraphExec_t exec = nullptr;
    int rc = HFH_OK;
    if (use_graph && !timing && R.exec && R.exec_x == R.x && R.exec_chunk == chunk) {
        exec = R.exec;
    } else if (use_graph && !timing) {
        if (R.exec) (void)hipGraphExecDestroy(R.exec);
        R.exec = nullptr;
        hipGraph_t graph = nullptr;
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
            for (int k = 0; k < chunk && !rc; ++k) rc = R.iteration();
            const hipError_t e2 = hipStreamEndCapture(s, &graph);
            if (rc || e2 != hipSuccess || hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
                (void)hipGetLastError();
                if (exec) (void)hipGraphExecDestroy(exec);
                exec = nullptr;
                rc = HFH_OK;  // fall back to direct launches
            }
            if (graph) (void)hipGraphDestroy(graph);
        } else {
            (void)hipGetLastError();
        }
        R.exec = exec;
        R.exec_x = R.x;
        R.exec_chunk = chunk;
    }
    // Chunk k+1 is launched before the host reads chunk k's state (copied
    // into one of two pinned slots, each with an event), unless chunk k is
    // expected to end the solve: the residual after it, extrapolated from
    // the last two chunks read, meets the tolerance.  A chunk launched after
    // the end finds the state done and every kernel returns at once, so the
    // iterates are those of the one-chunk-at-a-time loop.  (With timing on,
    // one chunk at a time.)
    struct Pipe {
        hipEvent_t ev[2] = {nullptr, nullptr};
        ~Pipe() {
            for (hipEvent_t e : ev)
                if (e) (void)hipEventDestroy(e);
        }
    } pipe;
    for (hipEvent_t &e : pipe.ev) HFH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    static_assert(sizeof(KState) <= 128 * sizeof(double), "two state slots in the pinned result buffer");
    KState *slot[2] = {reinterpret_cast<KState *>(c->host_res), reinterpret_cast<KState *>(c->host_res + 128)};
    std::vector<hipEvent_t> evs;
    auto launch_chunk = [&](int k) -> int {
        if (exec) {
            HFH_HIP(hipGraphLaunch(exec, s));
        } else {
            int r2 = HFH_OK;
            for (int i = 0; i < chunk && !r2; ++i) {
                if (timing) {
                    hipEvent_t a, b;
                    HFH_HIP(hipEventCreate(&a));
                    HFH_HIP(hipEventCreate(&b));
                    evs.push_back(a);
                    evs.push_back(b);
                    HFH_HIP(hipEventRecord(a, s));
                    r2 = R.spmv_kernel();
                    HFH_HIP(hipEventRecord(b, s));
                    if (!r2) r2 = R.fin_pap();
                    if (!r2) r2 = R.tail();
                } else {
                    r2 = R.iteration();
                }
            }
            if (r2) return r2;
        }
        HFH_HIP(hipMemcpyAsync(slot[k & 1], R.st, sizeof(KState), hipMemcpyDeviceToHost, s));
        HFH_HIP(hipEventRecord(pipe.ev[k & 1], s));
        return HFH_OK;
    };
    double ra = -1, rb = -1;   // residuals after the last two chunks read (ra the later)
    int queued = 0;            // chunks enqueued
    for (int k = 0;; ++k) {
        if (queued == k) {
            if ((rc = launch_chunk(k))) break;
            queued = k + 1;
        }
        if (!timing && static_cast<long long>(k + 1) * chunk <= maxits) {
            const KState &h = *slot[(k + 1) & 1];   // chunk k-1's copy: tolerances and res0
            const bool ends = k >= 2 && ra > 0 && rb > 0 &&
                              (ra * (ra / rb) < h.atol || ra * (ra / rb) < h.rtol * h.res0);
            if (!ends) {
                if ((rc = launch_chunk(k + 1))) break;
                queued = k + 2;
            }
        }
        HFH_HIP(hipEventSynchronize(pipe.ev[k & 1]));
        const KState h = *slot[k & 1];
        if (h.done) break;
        rb = ra;
        ra = h.res;
        if (static_cast<long long>(k + 1) * chunk > static_cast<long long>(maxits) + chunk) break;  // the device test stops at maxits
    }
    if (queued > 0) HFH_HIP(hipStreamSynchronize(s));
    if (!evs.empty()) {
        int it = 0;
        (void)hipStreamSynchronize(s);
        (void)hipMemcpy(&it, &R.st->iter, sizeof(int), hipMemcpyDeviceToHost);
        for (size_t k = 0; k + 1 < evs.size(); k += 2) {
            float ms = 0;
            if (static_cast<long long>(k / 2) < it && hipEventElapsedTime(&ms, evs[k], evs[k + 1]) == hipSuccess) {
                c->timing.ms[0] += ms;
                c->timing.launches[0] += 1;
            }
            (void)hipEventDestroy(evs[k]);
            (void)hipEventDestroy(evs[k + 1]);
        }
    }
    return rc;
}

// Workspace + state of a PCG solve of size n in type T.
template <typename T>
struct PcgWork {
    Work w;
    PcgRun<T> R;
    KState *st = nullptr;
    double *hist = nullptr;
    int maxits = 0;
    // search direction kept for a continued solve: p of the last iteration and
    // the r.z that formed its step
    bool have_dir = false;
    double prev_rz = 0;
    ~PcgWork() {
        if (R.exec) (void)hipGraphExecDestroy(R.exec);
    }

    // Buffers carved from `base` (cap bytes, 256-byte aligned) instead of
    // allocated: r, p, z, Ap, the partials, the state and (with_hist) the
    // residual history.  Jacobi or no preconditioner only.  Returns
    // HFH_ERR_INVALID when they do not fit.
    int init_static(const hfh_csr *A, hfh_prec *P, int maxits_, bool with_hist, char *base, size_t cap) {
        hfh_ctx *c = A->ctx;
        const int n = A->nrow;
        maxits = maxits_;
        R.A = A;
        R.P = P;
        R.n = n;
        R.mode = (!P || P->kind == HFH_PREC_NONE) ? 0 : 1;
        R.invd = R.mode == 1 ? static_cast<const T *>(P->invd) : nullptr;
        R.nb = stream_blocks(c, n);
        R.stride = kRedBlocks;
        auto up256 = [](size_t b) { return (b + 255) / 256 * 256; };
        const size_t vb = up256(sizeof(T) * static_cast<size_t>(std::max(n, 1)));
        const size_t bpart = up256(sizeof(double) * 2 * R.stride);
        const size_t bspart = up256(sizeof(double) * std::max(A->nwg, 1));
        const size_t bst = up256(sizeof(KState));
        const size_t bhist = with_hist ? up256(sizeof(double) * (static_cast<size_t>(maxits) + 2)) : 0;
        const size_t need = 4 * vb + bpart + bspart + bst + bhist;
        HFH_CHECK(need <= cap, HFH_ERR_INVALID,
                  "hfh_pcg_solve_noalloc: the solve needs %zu bytes of scratch (n = %d, %s%s), more than the "
                  "static %zu", need, n, sizeof(T) == 8 ? "FP64" : "FP32", with_hist ? ", with history" : "", cap);
        char *o = base;
        R.r = reinterpret_cast<T *>(o); o += vb;
        R.p = reinterpret_cast<T *>(o); o += vb;
        R.z = reinterpret_cast<T *>(o); o += vb;
        R.Ap = reinterpret_cast<T *>(o); o += vb;
        R.part = reinterpret_cast<double *>(o); o += bpart;
        R.spart = reinterpret_cast<double *>(o); o += bspart;
        st = reinterpret_cast<KState *>(o); o += bst;
        hist = with_hist ? reinterpret_cast<double *>(o) : nullptr;
        R.st = st;
        R.hist = hist;
        return HFH_OK;
    }

    int init(const hfh_csr *A, hfh_prec *P, int maxits_) {
        hfh_ctx *c = A->ctx;
        const int n = A->nrow;
        maxits = maxits_;
        R.A = A;
        R.P = P;
        R.n = n;
        R.mode = (!P || P->kind == HFH_PREC_NONE) ? 0 : (P->kind == HFH_PREC_JACOBI ? 1 : 2);
        R.invd = (P && P->kind == HFH_PREC_JACOBI) ? static_cast<const T *>(P->invd) : nullptr;
        R.f0 = R.mode == 2 && P->kind == HFH_PREC_MCSGS && P->ncolors > 1 && P->color_off[1] > 0;
        R.idv0 = R.f0 ? static_cast<const T *>(P->invd) : nullptr;
        R.n0 = R.f0 ? P->color_off[1] : 0;
        R.nb = stream_blocks(c, n);
        R.stride = kRedBlocks;
        int rc;
        const size_t vb = sizeof(T) * static_cast<size_t>(std::max(n, 1));
        if ((rc = w.alloc(reinterpret_cast<void **>(&R.r), vb)) || (rc = w.alloc(reinterpret_cast<void **>(&R.p), vb)) ||
            (rc = w.alloc(reinterpret_cast<void **>(&R.z), vb)) || (rc = w.alloc(reinterpret_cast<void **>(&R.Ap), vb)) ||
            (rc = w.alloc(reinterpret_cast<void **>(&R.part), sizeof(double) * 2 * R.stride)) ||
            (rc = w.alloc(reinterpret_cast<void **>(&R.spart), sizeof(double) * std::max(A->nwg, 1))) ||
            (rc = w.alloc(reinterpret_cast<void **>(&st), sizeof(KState))) ||
            (rc = w.alloc(reinterpret_cast<void **>(&hist), sizeof(double) * (static_cast<size_t>(maxits) + 2))))
            return rc;
        if (R.mode == 2 && P->kind == HFH_PREC_MCSGS &&
            (rc = w.alloc(reinterpret_cast<void **>(&R.rzp), sizeof(double) * mcsgs_rz_slots(P))))
            return rc;
        R.st = st;
        R.hist = hist;
        return HFH_OK;
    }

    // Solves A x = b from the given x (device pointers); fills info.
    // x_zero: x holds zeros, so r = b without the SpMV.  keep_direction:
    // continue the previous solve's search direction, p = z + (r.z / r_k.z_k) p
    // (the residual replacement of reliable-update CG), instead of p = z.
    // noalloc: no graph, no events, no pageable copies (the state travels as
    // kernel arguments and comes back through the context's pinned slots),
    // so nothing is allocated; time_ms is host wall time.
    int solve(const T *b, T *x, const hfh_control &ctl, hfh_solve_info *info, double *hhist, bool use_graph,
              bool x_zero = false, bool keep_direction = false, bool noalloc = false) {
        const hfh_csr *A = R.A;
        hfh_ctx *c = A->ctx;
        hipStream_t s = c->stream;
        const int n = R.n;
        R.x = x;
        {
            const uintptr_t al = reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(R.r) |
                                 reinterpret_cast<uintptr_t>(R.p) | reinterpret_cast<uintptr_t>(R.z) |
                                 reinterpret_cast<uintptr_t>(R.Ap) | reinterpret_cast<uintptr_t>(R.invd) |
                                 reinterpret_cast<uintptr_t>(R.idv0);
            R.nv = (al & 15) == 0 ? n / V16<T>::W : 0;
        }
        hfh_vec vr{c, n, A->dtype, R.r}, vb{c, n, A->dtype, const_cast<T *>(b)}, vz{c, n, A->dtype, R.z},
            vp{c, n, A->dtype, R.p};
        hipEvent_t e0 = nullptr, e1 = nullptr;
        const auto t0 = std::chrono::steady_clock::now();
        if (!noalloc) {
            HFH_HIP(hipEventCreate(&e0));
            HFH_HIP(hipEventCreate(&e1));
            HFH_HIP(hipEventRecord(e0, s));
        }
        int rc;
        if (x_zero) {
            HFH_HIP(hipMemcpyAsync(R.r, b, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
            rc = HFH_OK;
        } else {
            rc = spmv_launch_ex(A, x, R.r, 0, nullptr, nullptr, nullptr);
            if (!rc) rc = hfh_scaleadd(&vr, -1.0, &vb);
        }
        double rr = 0;
        if (!rc) rc = hfh_dot(&vr, &vr, &rr);
        if (rc) return rc;
        const T rrT = static_cast<T>(rr);
        const double res0 = static_cast<double>(static_cast<T>(sqrt(static_cast<double>(rrT))));
        const int state0 = host_check(0, res0, res0, ctl);
        double beta = rr;
        const bool cont = keep_direction && have_dir;
        if (R.mode != 0) {
            HFH_HIP(hipMemsetAsync(R.z, 0, sizeof(T) * n, s));
            if ((rc = prec_apply_raw(R.P, R.r, R.z))) return rc;
            if ((rc = hfh_dot(&vr, &vz, &beta))) return rc;
        }
        if (cont) {
            const T ratio = static_cast<T>(beta) / static_cast<T>(prev_rz);
            if ((rc = hfh_scaleadd(&vp, static_cast<double>(ratio), R.mode == 0 ? &vr : &vz))) return rc;
        } else {
            HFH_HIP(hipMemcpyAsync(R.p, R.mode == 0 ? R.r : R.z, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
        }
        KState h{};
        h.beta = static_cast<double>(static_cast<T>(beta));
        h.res = res0;
        h.res0 = res0;
        h.atol = ctl.atol;
        h.rtol = ctl.rtol;
        h.dtol = ctl.dtol;
        h.maxits = ctl.maxits;
        h.state = state0;
        h.done = state0 != HFH_ITERATE;
        h.idle = h.done;
        if (noalloc) {
            hipLaunchKernelGGL(k_set_state, dim3(1), dim3(64), 0, s, st, h, hist, res0);
            if ((rc = launch_check("k_set_state"))) return rc;
        } else {
            HFH_HIP(hipMemcpyAsync(st, &h, sizeof(KState), hipMemcpyHostToDevice, s));
            HFH_HIP(hipMemcpyAsync(hist, &res0, sizeof(double), hipMemcpyHostToDevice, s));
        }
        if (!h.done && (rc = pcg_loop(R, ctl.maxits, use_graph && !noalloc, !noalloc))) return rc;
        float ms = 0;
        if (noalloc) {
            static_assert(sizeof(KState) <= 256 * sizeof(double), "the pinned result slots hold the state");
            HFH_HIP(hipMemcpyAsync(c->host_res, st, sizeof(KState), hipMemcpyDeviceToHost, s));
            HFH_HIP(hipStreamSynchronize(s));
            std::memcpy(&h, c->host_res, sizeof(KState));
            ms = static_cast<float>(
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        } else {
            HFH_HIP(hipEventRecord(e1, s));
            HFH_HIP(hipMemcpyAsync(&h, st, sizeof(KState), hipMemcpyDeviceToHost, s));
            HFH_HIP(hipStreamSynchronize(s));
            (void)hipEventElapsedTime(&ms, e0, e1);
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        if (h.iter > 0) {
            have_dir = true;
            prev_rz = h.beta;
        }
        if (hhist && h.iter >= 0) {
            if (noalloc) {
                // through the context's 256 pinned result slots, a piece at a time
                for (int k = 0; k <= h.iter; k += 256) {
                    const int m = std::min(256, h.iter + 1 - k);
                    HFH_HIP(hipMemcpyAsync(c->host_res, hist + k, sizeof(double) * m, hipMemcpyDeviceToHost, s));
                    HFH_HIP(hipStreamSynchronize(s));
                    std::memcpy(hhist + k, c->host_res, sizeof(double) * m);
                }
            } else {
                HFH_HIP(hipMemcpy(hhist, hist, sizeof(double) * (h.iter + 1), hipMemcpyDeviceToHost));
            }
        }
        if (info) {
            info->iters = h.iter;
            info->state = h.state;
            info->res = h.res;
            info->res0 = h.res0;
            info->time_ms = ms;
        }
        return HFH_OK;
    }
};

// An empty system (n = 0) is solved exactly by the empty vector: no
// iteration, residual 0, reported as an absolute success.
static int empty_system(hfh_solve_info *info) {
    if (info) *info = hfh_solve_info{0, HFH_SUCCESS_ABS, 0.0, 0.0, 0.0};
    return HFH_OK;
}

static int solve_status(int state) {
    if (state == HFH_SUCCESS_ABS || state == HFH_SUCCESS_REL) return HFH_OK;
    set_error("solver stopped without convergence (state %d)", state);
    return HFH_ERR_NOT_CONVERGED;
}

static int check_solve_args(const hfh_csr *A, const hfh_prec *P, const hfh_vec *b, const hfh_vec *x,
                            const hfh_control *ctl, const char *what) {
    HFH_CHECK(A && b && x && ctl, HFH_ERR_INVALID, "%s: null argument", what);
    HFH_CHECK(A->has_structure && A->nrow == A->ncol, HFH_ERR_INVALID, "%s: matrix must be square with a structure", what);
    HFH_CHECK(b->n == A->nrow && x->n == A->nrow && b->dtype == A->dtype && x->dtype == A->dtype, HFH_ERR_INVALID,
              "%s: vector size/dtype mismatch", what);
    HFH_CHECK(!P || (P->n == A->nrow && P->dtype == A->dtype), HFH_ERR_INVALID, "%s: preconditioner mismatch", what);
    HFH_CHECK(ctl->maxits > 0 && ctl->atol >= 0 && ctl->rtol >= 0 && ctl->dtol > 1, HFH_ERR_INVALID,
              "%s: control needs maxits > 0, atol >= 0, rtol >= 0, dtol > 1", what);
    HFH_CHECK(x->d != b->d, HFH_ERR_INVALID, "%s: x and b must not alias", what);
    return HFH_OK;
}

}  // namespace hfh

using namespace hfh;

extern "C" int hfh_pcg_solve(const hfh_csr *A, hfh_prec *P, const hfh_vec *b, hfh_vec *x, const hfh_control *ctl,
                             hfh_solve_info *info, double *hist) {
    int rc = check_solve_args(A, P, b, x, ctl, "hfh_pcg_solve");
    if (rc) return rc;
    if (A->nrow == 0) return empty_system(info);
    if ((rc = prec_sync(P))) return rc;   // before the iterations are captured
    hfh_solve_info inf{};
    if (A->dtype == HFH_F64) {
        PcgWork<double> W;
        if ((rc = W.init(A, P, ctl->maxits))) return rc;
        if ((rc = W.solve((const double *)b->d, (double *)x->d, *ctl, &inf, hist, true))) return rc;
    } else {
        PcgWork<float> W;
        if ((rc = W.init(A, P, ctl->maxits))) return rc;
        if ((rc = W.solve((const float *)b->d, (float *)x->d, *ctl, &inf, hist, true))) return rc;
    }
    if ((rc = sweep_failed(P))) return rc;
    if (info) *info = inf;
    return solve_status(inf.state);
}

// ============================================================================
// Allocation-free CG for callers that cannot allocate device memory (code
// under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
// Its work vectors, partials and state live in `scratch`, a static device
// array of kStaticScratch bytes that every device the library runs on holds
// until the process ends.  Host code never uses the array's name as an
// address: hfh_ctx_create looks up each device's copy once
// (hipGetSymbolAddress, with that device current) and keeps it in the
// context.  All calls on one device share the array, so calls that overlap on
// one device race; every call writes what it reads before reading it.

static __device__ __align__(256) char scratch[kStaticScratch];

namespace hfh {

int static_scratch(int device, void **ptr) {
    static std::mutex mu;
    static void *addr[64] = {nullptr};
    HFH_CHECK(device >= 0 && device < 64, HFH_ERR_INVALID, "static scratch: device %d out of range", device);
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        HFH_CHECK(p != nullptr, HFH_ERR_HIP, "hipGetSymbolAddress(scratch): null address");
        addr[device] = p;
    }
    *ptr = addr[device];
    return HFH_OK;
}

}  // namespace hfh

extern "C" int hfh_pcg_solve_noalloc(const hfh_csr *A, hfh_prec *P, const hfh_vec *b, hfh_vec *x,
                                     const hfh_control *ctl, hfh_solve_info *info, double *hist) {
    int rc = check_solve_args(A, P, b, x, ctl, "hfh_pcg_solve_noalloc");
    if (rc) return rc;
    HFH_CHECK(!P || P->kind == HFH_PREC_NONE || P->kind == HFH_PREC_JACOBI, HFH_ERR_UNSUPPORTED,
              "hfh_pcg_solve_noalloc: Jacobi or no preconditioner only");
    HFH_CHECK(A->ctx->scratch, HFH_ERR_INVALID, "hfh_pcg_solve_noalloc: context has no static scratch");
    if (A->nrow == 0) return empty_system(info);
    char *base = static_cast<char *>(A->ctx->scratch);
    hfh_solve_info inf{};
    if (A->dtype == HFH_F64) {
        PcgWork<double> W;
        if ((rc = W.init_static(A, P, ctl->maxits, hist != nullptr, base, kStaticScratch))) return rc;
        if ((rc = W.solve((const double *)b->d, (double *)x->d, *ctl, &inf, hist, false, false, false, true)))
            return rc;
    } else {
        PcgWork<float> W;
        if ((rc = W.init_static(A, P, ctl->maxits, hist != nullptr, base, kStaticScratch))) return rc;
        if ((rc = W.solve((const float *)b->d, (float *)x->d, *ctl, &inf, hist, false, false, false, true)))
            return rc;
    }
    if (info) *info = inf;
    return solve_status(inf.state);
}

// ============================================================================
// GMRES / FGMRES (gmres.cc:285-510, fgmres.cc:75-262)
//
// Device: the Krylov basis, the preconditioned vectors, A z and the modified
// Gram-Schmidt sweep.  MGS step i fuses the previous axpy with the next dot
// (w -= h_{i-1} v_{i-1}; h_i = w.v_i): every workgroup of step i sums step
// i-1's partials itself, in the one-workgroup finisher's order (the same
// bits), while its first loads of w and v are in flight, so the host issues
// j+3 launches per step (j+2 MGS steps, the normalisation) and reads back one
// Hessenberg column.  Host: the Givens rotations, the residual estimate
// |g_{j+1}| and the back substitution, in the reference's order.

namespace hfh {

// Elements per thread loaded before the workgroup knows h: their loads
// overlap the partial sums' latency.
constexpr int kMgsPre = 8;

// With vprev: h_{i-1} = the previous step's partials pprev[0..np) summed as
// k_fin_mgs sums them, rounded to T (workgroup 0 stores it at hout_prev),
// and w -= h_{i-1} vprev.  Then the partials of w.vcur (w.w without vcur)
// into part.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mgs(int n, T *__restrict__ w, const T *__restrict__ vprev,
                                                const T *__restrict__ vcur, const double *__restrict__ pprev, int np,
                                                double *hout_prev, double *part) {
    __shared__ double sh[kFinBlock / 64];
    __shared__ double bc;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    const long long i0 = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x;
    T wp[kMgsPre], vp[kMgsPre], cp[kMgsPre];
#pragma unroll
    for (int k = 0; k < kMgsPre; ++k) {
        const long long i = i0 + k * stride;
        const bool in = i < n;
        wp[k] = in ? w[i] : static_cast<T>(0);
        vp[k] = (vprev && in) ? vprev[i] : static_cast<T>(0);
        cp[k] = (vcur && in) ? vcur[i] : static_cast<T>(0);
    }
    T h = 0;
    if (vprev) {
        const T hp = static_cast<T>(fin_sum_all(pprev, np, sh, &bc));
        if (blockIdx.x == 0 && threadIdx.x == 0) *hout_prev = static_cast<double>(hp);
        h = static_cast<T>(-1) * hp;
    }
    double s = 0;
#pragma unroll
    for (int k = 0; k < kMgsPre; ++k) {
        const long long i = i0 + k * stride;
        if (i < n) {
            T wi = wp[k];
            if (vprev) {
                wi += h * vp[k];
                w[i] = wi;
            }
            const T other = vcur ? cp[k] : wi;
            s += static_cast<double>(wi) * static_cast<double>(other);
        }
    }
    for (long long i = i0 + kMgsPre * stride; i < n; i += stride) {
        T wi = w[i];
        if (vprev) {
            wi += h * vprev[i];
            w[i] = wi;
        }
        const T other = vcur ? vcur[i] : wi;
        s += static_cast<double>(wi) * static_cast<double>(other);
    }
    const double v[1] = {block_sum(s, sh)};
    store_partials<1>(v, part, kRedBlocks, blockIdx.x);
}

// h = w.v (norm == false) or ||w|| (norm == true), rounded to T
template <typename T>
__global__ __launch_bounds__(kFinBlock) void k_fin_mgs(const double *part, int nb, bool norm, double *hout) {
    double out[1];
    reduce_partials<1>(part, kRedBlocks, nb, out);
    if (threadIdx.x == 0) {
        const T d = static_cast<T>(out[0]);
        *hout = norm ? static_cast<double>(static_cast<T>(sqrt(static_cast<double>(d)))) : static_cast<double>(d);
    }
}

// v = w * (1 / h)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale_into(int n, T *__restrict__ v, const T *__restrict__ w,
                                                       const double *h) {
    const T s = static_cast<T>(1) / static_cast<T>(*h);
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        v[i] = w[i] * s;
}

// v = w * (1 / h) with h = ||w|| from the last MGS step's partials of w.w
// (k_fin_mgs's arithmetic, summed in every workgroup while its first loads
// of w are in flight; workgroup 0 stores h at hout)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_norm_scale_into(int n, T *__restrict__ v, const T *__restrict__ w,
                                                            const double *__restrict__ pw, int np, double *hout) {
    __shared__ double sh[kFinBlock / 64];
    __shared__ double bc;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    const long long i0 = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x;
    T wp[kMgsPre];
#pragma unroll
    for (int k = 0; k < kMgsPre; ++k) wp[k] = i0 + k * stride < n ? w[i0 + k * stride] : static_cast<T>(0);
    const T d = static_cast<T>(fin_sum_all(pw, np, sh, &bc));
    const T h = static_cast<T>(sqrt(static_cast<double>(d)));
    if (blockIdx.x == 0 && threadIdx.x == 0) *hout = static_cast<double>(h);
    const T s = static_cast<T>(1) / h;
#pragma unroll
    for (int k = 0; k < kMgsPre; ++k)
        if (i0 + k * stride < n) v[i0 + k * stride] = wp[k] * s;
    for (long long i = i0 + kMgsPre * stride; i < n; i += stride) v[i] = w[i] * s;
}

constexpr int kMaxBasis = 256;

// The whole MGS sweep of one Arnoldi column in one cooperative launch
// (GMRES::SolveRight's inner loop, gmres.cc:285-510): h_i = w.v_i and
// w -= h_i v_i for i = 0..j, then h_{j+1} = ||w|| and v_{j+1} = w / h_{j+1},
// with the same per-element arithmetic as k_mgs and k_norm_scale_into.  Each
// thread keeps its E elements of w in registers for the whole sweep and
// loads the next basis vector while a step's sum travels, so a step reads
// only v_i from memory (k_mgs reads w, v_(i-1) and v_i and writes w).  The
// step's sum crosses the grid without a fence: every workgroup stores its
// partial with an agent-scope relaxed atomic store over a signalling-NaN
// sentinel, then polls all G partials (each value its own flag) and sums
// them in one fixed order, so every workgroup holds the same h_i.  The
// partials of the launches alternate between two halves of `slots` (par);
// each workgroup sets its own slots of the other half back to the sentinel
// for the next launch.  The launch is cooperative: all G workgroups are
// resident together.  Polls are bounded (kCoopSpins); a workgroup that gives
// up raises *fail and continues with the NaN.
constexpr int kCoopBlock = 1024;
constexpr int kCoopSpins = 1 << 16;
constexpr unsigned long long kCoopSentinel = 0x7FF4DEAD5EEDBEEFull;   // exponent all ones, quiet bit clear

// Sum of one step over the grid, valid in every thread: this workgroup's
// partial (wave trees, then the waves in order) published in slot row[g],
// then the G partials polled and summed (thread t holds row[t]; wave trees,
// then the waves in order).
__device__ __forceinline__ double coop_step_sum(double v, unsigned long long *row, double *sh, double *bc,
                                                int *fail) {
    const int G = gridDim.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    v = wave_sum(v);
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    if (tid == 0) {
        double t = 0;
        for (int k = 0; k < kCoopBlock / 64; ++k) t += sh[k];
        __hip_atomic_store(row + blockIdx.x, __builtin_bit_cast(unsigned long long, t), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    double x = 0;
    if (tid < G) {
        unsigned long long u = __hip_atomic_load(row + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int spins = 0;
        while (u == kCoopSentinel && spins < kCoopSpins) {
            ++spins;
            __builtin_amdgcn_s_sleep(1);
            u = __hip_atomic_load(row + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (u == kCoopSentinel) __hip_atomic_store(fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = __builtin_bit_cast(double, u);
    }
    __syncthreads();   // sh is reused
    x = wave_sum(x);
    if (lane == 0) sh[wv] = x;
    __syncthreads();
    if (tid == 0) {
        double t = 0;
        for (int k = 0; k < (G + 63) / 64; ++k) t += sh[k];
        *bc = t;
    }
    __syncthreads();
    return *bc;
}

template <typename T, int E>
__global__ __launch_bounds__(kCoopBlock) void k_mgs_coop(int n, const T *__restrict__ w, T *const *__restrict__ V,
                                                         int j, unsigned long long *slots, int nstep, int par,
                                                         double *Hd, int *fail) {
    __shared__ double sh[kCoopBlock / 64];
    __shared__ double bc;
    const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
    const long long span = static_cast<long long>(G) * kCoopBlock;
    const long long i0 = static_cast<long long>(g) * kCoopBlock + tid;
    if (tid < nstep) slots[(static_cast<long long>(1 - par) * nstep + tid) * G + g] = kCoopSentinel;
    T wr[E], vr[E], vn[E];
    {
        const T *v0 = V[0];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const long long i = i0 + e * span;
            wr[e] = i < n ? w[i] : static_cast<T>(0);
            vr[e] = i < n ? v0[i] : static_cast<T>(0);
        }
    }
    for (int i = 0; i <= j + 1; ++i) {   // step j + 1: the norm
        double sum = 0;
#pragma unroll
        for (int e = 0; e < E; ++e)
            sum += static_cast<double>(wr[e]) * static_cast<double>(i <= j ? vr[e] : wr[e]);
        if (i + 1 <= j) {   // the next basis vector, loaded while the sum travels
            const T *vnext = V[i + 1];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const long long k = i0 + e * span;
                vn[e] = k < n ? vnext[k] : static_cast<T>(0);
            }
        }
        const double tot = coop_step_sum(sum, slots + (static_cast<long long>(par) * nstep + i) * G, sh, &bc, fail);
        if (i <= j) {
            const T hi = static_cast<T>(tot);
            if (g == 0 && tid == 0) Hd[i] = static_cast<double>(hi);
            const T mh = static_cast<T>(-1) * hi;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                wr[e] += mh * vr[e];
                vr[e] = vn[e];
            }
        } else {
            const T d = static_cast<T>(tot);
            const T hn = static_cast<T>(sqrt(static_cast<double>(d)));
            if (g == 0 && tid == 0) Hd[j + 1] = static_cast<double>(hn);
            const T sc = static_cast<T>(1) / hn;
            T *vj1 = V[j + 1];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const long long k = i0 + e * span;
                if (k < n) vj1[k] = wr[e] * sc;
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_fill_u64(long long m, unsigned long long *p, unsigned long long v) {
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < m; i += (long long)gridDim.x * kBlock)
        p[i] = v;
}

// Workgroups of a cooperative MGS launch for n rows (one per CU, all
// resident), or 0 when it does not apply: more than E elements per thread,
// or no cooperative launch on the device.
template <typename T, int E>
static int coop_grid(hfh_ctx *c, int n) {
    int dev = 0, coop = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mgs_coop<T, E>, kCoopBlock, 0) != hipSuccess ||
        per_cu < 1)
        return 0;
    const int G = std::max(c->num_cu, 1);
    if (G > kCoopBlock || static_cast<long long>(G) * kCoopBlock * E < n) return 0;
    return G;
}

struct Coeffs {
    double y[kMaxBasis];
};

// x += sum_{j<=k} y_j V_j, added one basis vector after the other
template <typename T>
__global__ __launch_bounds__(kBlock) void k_multi_axpy(int n, T *__restrict__ x, T *const *__restrict__ V, int k1,
                                                       Coeffs c) {
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock) {
        T xi = x[i];
        for (int j = 0; j < k1; ++j) {
            // the basis pointers are device allocations: global loads, not flat
            const __attribute__((address_space(1))) T *vj = (const __attribute__((address_space(1))) T *)V[j];
            xi += static_cast<T>(c.y[j]) * vj[i];
        }
        x[i] = xi;
    }
}

// x += 1 * w  (Axpy(w, 1), gmres.cc:398)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_add_into(int n, T *__restrict__ x, const T *__restrict__ w) {
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        x[i] += static_cast<T>(1) * w[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_neg_add(int n, T *__restrict__ y, const T *__restrict__ b) {
    // y = -1 * y + b  (ScaleAdd(b, -1))
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        y[i] = static_cast<T>(-1) * y[i] + b[i];
}

template <typename T>
static int gmres_impl(const hfh_csr *A, hfh_prec *P, int size_basis, int flexible, const T *b, T *x,
                      const hfh_control &ctl, hfh_solve_info *info, double *hhist) {
    hfh_ctx *c = A->ctx;
    hipStream_t s = c->stream;
    const int n = A->nrow;
    const int m = std::min(size_basis, ctl.maxits);
    const int m1 = m + 1;
    const bool precond = P && P->kind != HFH_PREC_NONE;
    Work w;
    int rc;
    const size_t vb = sizeof(T) * static_cast<size_t>(std::max(n, 1));
    std::vector<T *> V(m1, nullptr), Z(m1, nullptr);
    for (int i = 0; i < m1; ++i)
        if ((rc = w.alloc(reinterpret_cast<void **>(&V[i]), vb))) return rc;
    if (flexible && precond)
        for (int i = 0; i < m1; ++i)
            if ((rc = w.alloc(reinterpret_cast<void **>(&Z[i]), vb))) return rc;
    T *wv = nullptr, *z = nullptr, **Vd = nullptr, **Zd = nullptr;
    double *part = nullptr, *Hd = nullptr;
    const int nb = stream_blocks(c, n);
    if ((rc = w.alloc(reinterpret_cast<void **>(&wv), vb)) || (rc = w.alloc(reinterpret_cast<void **>(&z), vb)) ||
        (rc = w.alloc(reinterpret_cast<void **>(&Vd), sizeof(T *) * m1)) ||
        (rc = w.alloc(reinterpret_cast<void **>(&Zd), sizeof(T *) * m1)) ||
        (rc = w.alloc(reinterpret_cast<void **>(&part), sizeof(double) * 2 * kRedBlocks)) ||
        (rc = w.alloc(reinterpret_cast<void **>(&Hd), sizeof(double) * (m1 + 2))))
        return rc;
    HFH_HIP(hipMemcpyAsync(Vd, V.data(), sizeof(T *) * m1, hipMemcpyHostToDevice, s));
    // The MGS sweep of a column in one cooperative launch where n fits, with
    // E elements per thread the smallest of 2, 4, 8 that covers n.  Below 2
    // per thread (n <= 1024 x CUs) a step is latency-bound either way, and
    // the launch per step was faster (upwind 40^3 GMRES(30) + ILU(0): 13.1
    // against 13.9 ms); at 8 (upwind 128^3) the sweep of a column takes
    // 113 us against about 240 us.
    int cg = 0, ce = 0;
    if (n > std::max(c->num_cu, 1) * kCoopBlock)
        for (int e : {2, 4, 8}) {
            const int G = e == 2 ? coop_grid<T, 2>(c, n) : e == 4 ? coop_grid<T, 4>(c, n) : coop_grid<T, 8>(c, n);
            if (G > 0) {
                cg = G;
                ce = e;
                break;
            }
        }
    (void)hipGetLastError();
    const int nstep = m1 + 1;   // steps of a column: j + 2 <= m + 1
    unsigned long long *slots = nullptr;
    int *cfail = nullptr;
    int cseq = 0;
    if (cg > 0) {
        const long long ns = 2LL * nstep * cg;
        if ((rc = w.alloc(reinterpret_cast<void **>(&slots), sizeof(unsigned long long) * ns)) ||
            (rc = w.alloc(reinterpret_cast<void **>(&cfail), sizeof(int))))
            return rc;
        hipLaunchKernelGGL(k_fill_u64, dim3(static_cast<unsigned>(std::min<long long>((ns + kBlock - 1) / kBlock, 1024))),
                           dim3(kBlock), 0, s, ns, slots, kCoopSentinel);
        if ((rc = launch_check("k_fill_u64"))) return rc;
        HFH_HIP(hipMemsetAsync(cfail, 0, sizeof(int), s));
    }
    auto mgs_coop = [&](int j) -> int {
        int par = cseq++ & 1;
        int nn = n, jj = j, ns = nstep;
        const T *wc = wv;
        T *const *Vc = Vd;
        void *args[] = {&nn, &wc, &Vc, &jj, &slots, &ns, &par, &Hd, &cfail};
        const void *f = ce == 2 ? (const void *)k_mgs_coop<T, 2>
                        : ce == 4 ? (const void *)k_mgs_coop<T, 4> : (const void *)k_mgs_coop<T, 8>;
        HFH_HIP(hipLaunchCooperativeKernel(f, dim3(cg), dim3(kCoopBlock), args, 0, s));
        return HFH_OK;
    };
    if (flexible && precond) HFH_HIP(hipMemcpyAsync(Zd, Z.data(), sizeof(T *) * m1, hipMemcpyHostToDevice, s));

    // pinned staging of the Hessenberg columns in two slots (the host reads
    // column j while the copy of column j+1 may be in flight), and their
    // events; released after the stream has drained
    struct Staging {
        hipStream_t s = nullptr;
        double *col = nullptr;
        hipEvent_t ev[2] = {nullptr, nullptr};
        ~Staging() {
            if (s) (void)hipStreamSynchronize(s);
            for (hipEvent_t e : ev)
                if (e) (void)hipEventDestroy(e);
            if (col) (void)hipHostFree(col);
        }
    } stg;
    stg.s = s;
    HFH_HIP(hipHostMalloc(reinterpret_cast<void **>(&stg.col), sizeof(double) * 2 * (m1 + 1)));
    for (hipEvent_t &e : stg.ev) HFH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));

    hipEvent_t e0, e1;
    HFH_HIP(hipEventCreate(&e0));
    HFH_HIP(hipEventCreate(&e1));
    HFH_HIP(hipEventRecord(e0, s));

    auto residual = [&](T *dst) -> int {
        int r2 = spmv_launch_ex(A, x, dst, 0, nullptr, nullptr, nullptr);
        if (r2) return r2;
        hipLaunchKernelGGL(k_neg_add<T>, dim3(nb), dim3(kBlock), 0, s, n, dst, b);
        return launch_check("k_neg_add");
    };
    // the MGS steps' partials alternate between two buffers: a step reads
    // the previous step's while it writes its own
    int ph = 0;
    auto pbuf = [&](int k) { return part + static_cast<size_t>(k & 1) * kRedBlocks; };
    // one MGS launch: with vprev, h_prev from the previous launch's partials
    // (stored at hprev_out) and w -= h_prev v_prev; partials of w.vcur or w.w
    auto mgs = [&](T *wv_, const T *vprev, const T *vcur, double *hprev_out) -> int {
        const double *pp = pbuf(ph);
        ++ph;
        hipLaunchKernelGGL(k_mgs<T>, dim3(nb), dim3(kBlock), 0, s, n, wv_, vprev, vcur, pp, nb, hprev_out, pbuf(ph));
        return launch_check("k_mgs");
    };
    auto norm = [&](T *v, double *out) -> int {
        int r2 = mgs(v, nullptr, nullptr, nullptr);
        if (r2) return r2;
        hipLaunchKernelGGL(k_fin_mgs<T>, dim3(1), dim3(kFinBlock), 0, s, (const double *)pbuf(ph), nb, true, Hd + m1);
        if ((r2 = launch_check("k_fin_mgs"))) return r2;
        return fetch(c, Hd + m1, out, 1);
    };
    // apply the preconditioner (identity without one) into dst; every
    // preconditioner's apply writes every entry of dst, so dst is not cleared
    auto precond_into = [&](const T *in, T *dst) -> int {
        if (!precond) {
            HFH_HIP(hipMemcpyAsync(dst, in, vb, hipMemcpyDeviceToDevice, s));
            return HFH_OK;
        }
        return prec_apply_raw(P, in, dst);
    };

    std::vector<double> H(static_cast<size_t>(m) * m1, 0.0), g(m1), cs(m1), sn(m1), col(m1 + 1);
    auto Hc = [&](int j, int i) -> double & { return H[static_cast<size_t>(j) * m1 + i]; };
    auto rot = [](double csv, double snv, double &dx, double &dy) {
        const T t = static_cast<T>(csv) * static_cast<T>(dx) + static_cast<T>(snv) * static_cast<T>(dy);
        dy = static_cast<T>(-snv) * static_cast<T>(dx) + static_cast<T>(csv) * static_cast<T>(dy);
        dx = t;
    };
    // x += M^-1 (V y) (GMRES) or x += Z y (FGMRES), y from the back substitution
    auto update = [&](int k) -> int {
        Coeffs cf;
        std::vector<double> y(g.begin(), g.end());
        for (int i = k + 1; i--;) {
            y[i] = static_cast<T>(y[i]) / static_cast<T>(Hc(i, i));
            const T t = static_cast<T>(y[i]);
            for (int j = 0; j < i; ++j) y[j] = static_cast<T>(y[j]) - static_cast<T>(Hc(i, j)) * t;
        }
        for (int j = 0; j <= k; ++j) cf.y[j] = y[j];
        if (flexible && precond) {
            hipLaunchKernelGGL(k_multi_axpy<T>, dim3(nb), dim3(kBlock), 0, s, n, x, Zd, k + 1, cf);
            return launch_check("k_multi_axpy");
        }
        HFH_HIP(hipMemsetAsync(z, 0, vb, s));
        hipLaunchKernelGGL(k_multi_axpy<T>, dim3(nb), dim3(kBlock), 0, s, n, z, Vd, k + 1, cf);
        int r2 = launch_check("k_multi_axpy");
        if (r2) return r2;
        if ((r2 = precond_into(z, wv))) return r2;
        hipLaunchKernelGGL(k_add_into<T>, dim3(nb), dim3(kBlock), 0, s, n, x, (const T *)wv);
        return launch_check("k_add_into");
    };

    int iter = 0;
    if ((rc = residual(V[0]))) return rc;
    double res = 0;
    if ((rc = norm(V[0], &res))) return rc;
    const double first = res;
    int state = host_check(0, res, first, ctl);
    std::vector<double> hist;
    hist.push_back(res);
    while (state == HFH_ITERATE) {
        std::fill(g.begin(), g.end(), 0.0);
        std::fill(H.begin(), H.end(), 0.0);
        {
            // V0 *= 1/res
            HFH_HIP(hipMemcpyAsync(Hd + m1 + 1, &res, sizeof(double), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_scale_into<T>, dim3(nb), dim3(kBlock), 0, s, n, V[0], (const T *)V[0], (const double *)(Hd + m1 + 1));
            if ((rc = launch_check("k_scale_into"))) return rc;
        }
        g[0] = res;
        // the device work of Arnoldi step j: z_j = M^-1 v_j, w = A z_j, the
        // MGS sweep and v_{j+1}; then the copy of H's column j to the host
        auto enqueue_step = [&](int j) -> int {
            T *zj = (flexible && precond) ? Z[j] : z;
            const T *az = zj;
            int r2;
            if (precond) {
                if ((r2 = precond_into(V[j], zj))) return r2;
            } else {
                az = V[j];
            }
            if ((r2 = spmv_launch_ex(A, az, wv, 0, nullptr, nullptr, nullptr))) return r2;
            if (cg > 0) {
                if ((r2 = mgs_coop(j))) return r2;
            } else {
                // step i stores h_{i-1}
                for (int i = 0; i <= j; ++i)
                    if ((r2 = mgs(wv, i > 0 ? V[i - 1] : nullptr, V[i], i > 0 ? Hd + i - 1 : nullptr))) return r2;
                // w -= h_j v_j (storing h_j); h_{j+1} = ||w||, v_{j+1} = w / h_{j+1}
                if ((r2 = mgs(wv, V[j], nullptr, Hd + j))) return r2;
                hipLaunchKernelGGL(k_norm_scale_into<T>, dim3(nb), dim3(kBlock), 0, s, n, V[j + 1], (const T *)wv,
                                   (const double *)pbuf(ph), nb, Hd + j + 1);
                if ((r2 = launch_check("k_norm_scale_into"))) return r2;
            }
            HFH_HIP(hipMemcpyAsync(stg.col + (j & 1) * (m1 + 1), Hd, sizeof(double) * (j + 2),
                                   hipMemcpyDeviceToHost, s));
            HFH_HIP(hipEventRecord(stg.ev[j & 1], s));
            return HFH_OK;
        };
        int j = 0, queued = 0;   // steps 0 .. queued-1 of this cycle are enqueued
        for (j = 0; j != m; ++j) {
            ++iter;
            if (queued == j) {
                if ((rc = enqueue_step(j))) return rc;
                queued = j + 1;
            }
            // Step j+1 is enqueued before step j's column is read, so the
            // device runs it while the host applies the rotations, unless
            // step j is expected to end the solve (its residual estimate,
            // extrapolated from the last two, meets the tolerance, or the
            // iteration limit is reached).  A step enqueued past the end
            // writes only v_{j+2}, z_{j+1}, w, z and H's next column, which
            // the update does not read.
            if (j + 1 < m && iter < ctl.maxits) {
                const size_t h = hist.size();
                const double est = h >= 2 && hist[h - 2] > 0 ? hist[h - 1] * (hist[h - 1] / hist[h - 2]) : res;
                if (!(est < ctl.atol || est < ctl.rtol * first)) {
                    if ((rc = enqueue_step(j + 1))) return rc;
                    queued = j + 2;
                }
            }
            HFH_HIP(hipEventSynchronize(stg.ev[j & 1]));
            for (int i = 0; i <= j + 1; ++i) col[i] = stg.col[(j & 1) * (m1 + 1) + i];
            for (int i = 0; i <= j + 1; ++i) Hc(j, i) = col[i];
            if (Hc(j, j + 1) == 0.0) {
                set_error("GMRES: breakdown, H(%d,%d) = 0", j, j + 1);
                return HFH_ERR_INVALID;
            }
            for (int k2 = 0; k2 < j; ++k2) rot(cs[k2], sn[k2], Hc(j, k2), Hc(j, k2 + 1));
            {
                const T dx = static_cast<T>(Hc(j, j)), dy = static_cast<T>(Hc(j, j + 1));
                const T beta = static_cast<T>(sqrt(static_cast<double>(dx * dx + dy * dy)));
                cs[j] = dx / beta;
                sn[j] = dy / beta;
            }
            rot(cs[j], sn[j], Hc(j, j), Hc(j, j + 1));
            rot(cs[j], sn[j], g[j], g[j + 1]);
            res = std::fabs(static_cast<T>(g[j + 1]));
            hist.push_back(res);
            state = host_check(iter, res, first, ctl);
            if (state != HFH_ITERATE) {
                if ((rc = update(j))) return rc;
                break;
            }
        }
        if (state == HFH_ITERATE) {
            if ((rc = update(m - 1))) return rc;
            if ((rc = residual(V[0]))) return rc;
        }
    }
    HFH_HIP(hipEventRecord(e1, s));
    HFH_HIP(hipStreamSynchronize(s));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (cfail) {
        int f = 0;
        HFH_HIP(hipMemcpy(&f, cfail, sizeof(int), hipMemcpyDeviceToHost));
        HFH_CHECK(!f, HFH_ERR_HIP, "GMRES: a cooperative MGS step gave up waiting for another workgroup's partial");
    }
    if (hhist) std::copy(hist.begin(), hist.end(), hhist);
    if (info) {
        info->iters = iter;
        info->state = state;
        info->res = res;
        info->res0 = first;
        info->time_ms = ms;
    }
    return HFH_OK;
}

// ============================================================================
// mixed-precision defect correction (mpir.cc:69-327)

template <typename TO, typename TI>
__global__ __launch_bounds__(kBlock) void k_cast_n(int n, TO *__restrict__ y, const TI *__restrict__ x) {
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        y[i] = static_cast<TO>(x[i]);
}

// x += (double) c  (CastInteriorFrom + Axpy(cor, 1), mpir.cc:223-253)
__global__ __launch_bounds__(kBlock) void k_add_correction(int n, double *__restrict__ x, const float *__restrict__ c) {
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        x[i] += 1.0 * static_cast<double>(c[i]);
}

}  // namespace hfh

extern "C" int hfh_gmres_solve(const hfh_csr *A, hfh_prec *P, int size_basis, int flexible, const hfh_vec *b,
                               hfh_vec *x, const hfh_control *ctl, hfh_solve_info *info, double *hist) {
    int rc = check_solve_args(A, P, b, x, ctl, "hfh_gmres_solve");
    if (rc) return rc;
    HFH_CHECK(size_basis > 0 && size_basis <= kMaxBasis, HFH_ERR_INVALID,
              "hfh_gmres_solve: basis size %d outside 1..%d", size_basis, kMaxBasis);
    if (A->nrow == 0) return empty_system(info);
    hfh_solve_info inf{};
    if (A->dtype == HFH_F64)
        rc = gmres_impl<double>(A, P, size_basis, flexible, (const double *)b->d, (double *)x->d, *ctl, &inf, hist);
    else
        rc = gmres_impl<float>(A, P, size_basis, flexible, (const float *)b->d, (float *)x->d, *ctl, &inf, hist);
    if (rc) return rc;
    if ((rc = sweep_failed(P))) return rc;
    if (info) *info = inf;
    return solve_status(inf.state);
}

extern "C" int hfh_dc_solve_ex(const hfh_csr *Ah, const hfh_csr *Al, hfh_prec *Pl, const hfh_vec *b, hfh_vec *x,
                               const hfh_control *outer, const hfh_control *inner, const hfh_dc_options *opt,
                               hfh_dc_info *info, double *hhist) {
    HFH_CHECK(Ah && Al && b && x && outer && inner, HFH_ERR_INVALID, "hfh_dc_solve: null argument");
    const bool keep = opt && opt->keep_direction;
    HFH_CHECK(Ah->dtype == HFH_F64 && Al->dtype == HFH_F32, HFH_ERR_INVALID,
              "hfh_dc_solve: needs an FP64 outer and an FP32 inner matrix");
    HFH_CHECK(Ah->nrow == Al->nrow && Ah->nnz == Al->nnz && Ah->nrow == Ah->ncol, HFH_ERR_INVALID,
              "hfh_dc_solve: matrices differ");
    HFH_CHECK(b->n == Ah->nrow && x->n == Ah->nrow && b->dtype == HFH_F64 && x->dtype == HFH_F64, HFH_ERR_INVALID,
              "hfh_dc_solve: vectors must be FP64 of size %d", Ah->nrow);
    HFH_CHECK(Ah->has_structure && Al->has_structure, HFH_ERR_INVALID, "hfh_dc_solve: matrices need a structure");
    HFH_CHECK(!Pl || (Pl->dtype == HFH_F32 && Pl->n == Al->nrow), HFH_ERR_INVALID, "hfh_dc_solve: bad preconditioner");
    HFH_CHECK(x->d != b->d, HFH_ERR_INVALID, "hfh_dc_solve: x and b must not alias");
    HFH_CHECK(outer->maxits > 0 && outer->atol >= 0 && outer->rtol >= 0 && outer->dtol > 1, HFH_ERR_INVALID,
              "hfh_dc_solve: outer control needs maxits > 0, atol >= 0, rtol >= 0, dtol > 1");
    HFH_CHECK(inner->maxits > 0 && inner->atol >= 0 && inner->rtol >= 0 && inner->dtol > 1, HFH_ERR_INVALID,
              "hfh_dc_solve: inner control needs maxits > 0, atol >= 0, rtol >= 0, dtol > 1");
    // the low precision tolerances are DataType_low members of the reference's
    // MPIR (mpir.h:225-240): its gates and the FP32 solver see them rounded
    const hfh_control inl{inner->maxits, (double)(float)inner->atol, (double)(float)inner->rtol,
                          (double)(float)inner->dtol};
    if (int rs = prec_sync(Pl)) return rs;   // before the inner iterations are captured
    if (Ah->nrow == 0) {
        // an empty system is solved exactly by the empty vector
        if (info) *info = hfh_dc_info{0, 0, HFH_SUCCESS_ABS, 0.0, 0.0, 0.0};
        return HFH_OK;
    }
    hfh_ctx *c = Ah->ctx;
    hipStream_t s = c->stream;
    const int n = Ah->nrow;
    Work w;
    double *rh = nullptr;
    float *rl = nullptr, *cl = nullptr;
    int rc;
    if ((rc = w.alloc(reinterpret_cast<void **>(&rh), sizeof(double) * std::max(n, 1))) ||
        (rc = w.alloc(reinterpret_cast<void **>(&rl), sizeof(float) * std::max(n, 1))) ||
        (rc = w.alloc(reinterpret_cast<void **>(&cl), sizeof(float) * std::max(n, 1))))
        return rc;
    PcgWork<float> W;
    if ((rc = W.init(Al, Pl, inner->maxits))) return rc;
    const int nb = stream_blocks(c, n);
    hfh_vec vr{c, n, HFH_F64, rh}, vb{c, n, HFH_F64, b->d};
    hipEvent_t e0, e1;
    HFH_HIP(hipEventCreate(&e0));
    HFH_HIP(hipEventCreate(&e1));
    HFH_HIP(hipEventRecord(e0, s));
    auto true_residual = [&](double *res) -> int {
        int r2 = spmv_launch_ex(Ah, x->d, rh, 0, nullptr, nullptr, nullptr);
        if (!r2) r2 = hfh_scaleadd(&vr, -1.0, &vb);
        if (!r2) r2 = hfh_nrm2(&vr, res);
        return r2;
    };
    double res = 0;
    if ((rc = true_residual(&res))) return rc;
    double res0 = res;
    const double first = res;
    std::vector<double> hist{res};
    int n_ecl = 0, inner_iters = 0;
    auto ecl_ok = [&]() {
        return n_ecl < outer->maxits && res > inl.atol && res > res0 * inl.rtol && res < inl.dtol;
    };
    while (res > outer->atol && res > res0 * outer->rtol && res < outer->dtol && ecl_ok()) {
        ++n_ecl;
        hipLaunchKernelGGL((k_cast_n<float, double>), dim3(nb), dim3(kBlock), 0, s, n, rl, (const double *)rh);
        if ((rc = launch_check("k_cast_n"))) return rc;
        HFH_HIP(hipMemsetAsync(cl, 0, sizeof(float) * n, s));
        hfh_solve_info ii{};
        if ((rc = W.solve(rl, cl, inl, &ii, nullptr, true, true, keep))) return rc;
        inner_iters += ii.iters;
        hipLaunchKernelGGL(k_add_correction, dim3(nb), dim3(kBlock), 0, s, n, (double *)x->d, (const float *)cl);
        if ((rc = launch_check("k_add_correction"))) return rc;
        if ((rc = true_residual(&res))) return rc;
        res0 = res;
        hist.push_back(res);
    }
    HFH_HIP(hipEventRecord(e1, s));
    HFH_HIP(hipStreamSynchronize(s));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int state;
    if (res < outer->atol) state = HFH_SUCCESS_ABS;
    else if (res < res0 * outer->rtol) state = HFH_SUCCESS_REL;
    else if (res >= outer->dtol) state = HFH_FAIL_DIV;
    else state = HFH_FAIL_MAXITS;
    if (hhist) std::copy(hist.begin(), hist.end(), hhist);
    if (info) {
        info->outer_iters = n_ecl;
        info->inner_iters = inner_iters;
        info->state = state;
        info->res = res;
        info->res0 = first;
        info->time_ms = ms;
    }
    return solve_status(state);
}

extern "C" int hfh_dc_solve(const hfh_csr *Ah, const hfh_csr *Al, hfh_prec *Pl, const hfh_vec *b, hfh_vec *x,
                            const hfh_control *outer, const hfh_control *inner, hfh_dc_info *info, double *hhist) {
    return hfh_dc_solve_ex(Ah, Al, Pl, b, x, outer, inner, nullptr, info, hhist);
}
