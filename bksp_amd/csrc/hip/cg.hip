// cg.hip -- conjugate gradients on the device (replaces src/cg.c:6-71).
//
// Plain FP64 path: two kernels per iteration (bkh_cgfused.hpp), K iterations and a
// one-wave tail captured in a hipGraph:
//   A: p = r + beta p folded into the SpMV's gather, q = A p, p.q partials
//      (cg.c:35-39, :53-54), after the convergence test of the previous iteration;
//   B: alpha = (r.r)/(p.q) from A's partials, x += alpha p, r -= alpha q, r.r
//      partials (cg.c:42-48).
// Matrices with rows longer than a wave block use an unfused sequence (SpMV with
// p.q, finalize, x/r update, finalize, p update). Scalars never leave the device;
// the host reads the control block once per graph replay.
//
// BKH_FP32_IR: mixed precision with reliable updates (Clark et al., "Solving
// lattice QCD systems of equations using mixed precision solvers on GPUs", 2010).
// The Krylov iteration runs in FP32 (values, 16-bit column deltas where they fit,
// vectors; FP64 scalars and dot accumulation) on a correction x32; each time its
// recursive residual has fallen by a factor delta (RU_DELTA unless bkh_opts.inner_tol
// sets it) the host replaces it by the FP64 true residual: x += x32 and r = b - A x in
// FP64, r32 = r, x32 = 0, and the FP32 iteration continues with its search
// direction, so no Krylov information is discarded (a restarted refinement would
// lose it at every pass). On SELL matrices the FP32 iteration is one kernel per
// iteration (k_cg1_sell: residual norm from the alpha recurrence, q = A p by
// recurrence so only r and q are gathered, x update deferred into the next kernel),
// elsewhere the two-kernel form of the FP64 path.
#include <math.h>
#include <chrono>
#include <mutex>
#include "bkh_cgfused.hpp"
#include "bkh_host.hpp"
#include "bkh_solver_common.hpp"
#include "bkh_spmv.hpp"

enum { CG_RR = 0, CG_ALPHA = 1, CG_BETA = 2 };

struct CgAlphaFin {
    __device__ void operator()(Ctl *c, const double *tot) const {
        double pq = tot[0];
        if (bad(pq)) { ctl_breakdown(c); return; }
        c->s[CG_ALPHA] = c->s[CG_RR] / pq;
    }
};

struct CgBetaFin {
    __device__ void operator()(Ctl *c, const double *tot) const {
        double rr = tot[0];
        c->s[CG_BETA] = rr / c->s[CG_RR];
        c->s[CG_RR] = rr;
        ctl_check(c, rr);
    }
};

// r = b - A x0 (q holds A x0), p = r; r.r and b.b  (cg.c:23-30)
struct CgInitFin {
    __device__ void operator()(Ctl *c, const double *tot) const {
        c->s[CG_RR] = tot[0];
        c->s[CG_BETA] = 0.0;
        c->fnb = sqrt(tot[1]);
        if (!(c->fnb > 0.0)) ctl_breakdown(c);
    }
};

#define ROWLOOP(i, n) \
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < (n); i += (long long)gridDim.x * 256)

// r = b - q; p_a = r (unfused path) or p_a = 0 (fused: beta = 0 gives p = r).
template <typename T>
__global__ __launch_bounds__(256) void k_cg_init(long long n, const T *b, const T *q, T *r, T *p,
                                                 int fused, double *part) {
    double v[2] = {0.0, 0.0};
    ROWLOOP(i, n) {
        T ri = b[i] - q[i];
        r[i] = ri;
        p[i] = fused ? (T)0 : ri;
        v[0] += (double)ri * (double)ri;
        v[1] += (double)b[i] * (double)b[i];
    }
    block_partial<2, 256>(v, part);
}

template <typename T>
__global__ __launch_bounds__(256) void k_cg_xr(long long n, T *x, T *r, const T *p, const T *q,
                                               Ctl *c, double *part) {
    if (ctl_stopped(c)) return;
    const T al = (T)c->s[CG_ALPHA];
    double v[1] = {0.0};
    ROWLOOP(i, n) {
        x[i] = fma(al, p[i], x[i]);
        T ri = fma(-al, q[i], r[i]);
        r[i] = ri;
        v[0] += (double)ri * (double)ri;
    }
    block_partial<1, 256>(v, part);
}

template <typename T>
__global__ __launch_bounds__(256) void k_cg_p(long long n, T *p, const T *r, const Ctl *c) {
    if (ctl_stopped(c)) return;
    const T be = (T)c->s[CG_BETA];
    ROWLOOP(i, n) p[i] = add_rn(mul_rn(be, p[i]), r[i]);
}

static inline bool cgf_sell(const bkh_matrix *A) {
    return A->nsl > 0 && (A->variant == SPMV_SELL || A->variant == SPMV_SELL32);
}

// The CG iteration kernels run on the slice-diagonal copy when the matrix has one and
// the default SELL-64 SpMV is selected (SELL32 keeps the int-column kernels).
static inline bool cg_dia(const bkh_matrix *A) {
    return A->dia_entries > 0 && A->nsl > 0 && A->variant == SPMV_SELL;
}

static inline bool cg_can_fuse(const bkh_matrix *A) { return cgf_sell(A) || A->max_row <= SPMV_WCAP; }

struct CgfGrid { int GA, chunk, GB, BSA; bool sell; };

static CgfGrid cgf_grid(const bkh_matrix *A) {
    CgfGrid g;
    g.sell = cgf_sell(A);
    g.BSA = 256;
    if (g.sell) {
        // 512 workgroups of 1024 threads once there are that many 16-slice chunks
        // (kernel B re-sums A's partials: 28.0 ms per FP64 128^3 solve against 29.4 ms
        // with 2048 workgroups of 256 threads; on the slice-diagonal copy 24.31 ms
        // against 24.55 ms with 1024 workgroups)
        if (A->nsl / 16 >= 512) {
            g.BSA = 1024;
            g.GA = 512;
        } else
            g.GA = std::max(1, std::min((A->nsl + 3) / 4, CGF_NPART));
        g.chunk = (A->nsl + g.GA - 1) / g.GA;
    } else {
        g.GA = std::max(1, std::min((A->nwblk + CGF_WPB - 1) / CGF_WPB, CGF_NPART));
        g.chunk = (A->nwblk + g.GA - 1) / g.GA;
    }
    // kernel B: 4096 rows per 512-thread workgroup (measured ~2 % faster per FP64 solve
    // than 2048: half the partials for every workgroup of kernel A to re-sum)
    g.GB = (int)std::max(1LL, std::min((long long)((A->n + 4095) / 4096), (long long)CGF_NPART));
    return g;
}

template <typename VT, typename T, int BS>
static void launch_cgf_a_sell(hipStream_t s, const bkh_matrix *A, const CgfGrid &g, T *r, T *P0, T *P1, T *q,
                              Ctl *c, int j, double *partB, double *partA) {
    const VT *sv;
    if constexpr (sizeof(VT) == sizeof(float)) sv = (const VT *)A->sval32;
    else sv = (const VT *)A->sval;
    const VT *dv;
    if constexpr (sizeof(VT) == sizeof(float)) dv = (const VT *)A->dval32;
    else dv = (const VT *)A->dval;
    if (cg_dia(A) && dv)
        hipLaunchKernelGGL((k_cgf_a_sell<VT, T, int, BS, true>), dim3(g.GA), dim3(BS), 0, s, (int)A->n, A->nsl,
                           g.chunk, A->rp, A->dso, dv, (const int *)A->doff, (const T *)r, P0, P1, q, c, j,
                           (const double *)partB, g.GB, partA);
    else if (A->scd && A->variant == SPMV_SELL)
        hipLaunchKernelGGL((k_cgf_a_sell<VT, T, short, BS>), dim3(g.GA), dim3(BS), 0, s, (int)A->n, A->nsl,
                           g.chunk, A->rp, A->sso, sv, (const short *)A->scd, (const T *)r, P0, P1, q, c, j,
                           (const double *)partB, g.GB, partA);
    else
        hipLaunchKernelGGL((k_cgf_a_sell<VT, T, int, BS>), dim3(g.GA), dim3(BS), 0, s, (int)A->n, A->nsl,
                           g.chunk, A->rp, A->sso, sv, (const int *)A->scol, (const T *)r, P0, P1, q, c, j,
                           (const double *)partB, g.GB, partA);
}

// Issues K fused iterations plus the tail on stream s, on vectors T with values VT
// (cd: 16-bit column deltas or null): the body of the iteration graph, and the
// iteration of the allocation-free entry point, which launches it directly.
template <typename VT, typename T>
static void cgf_issue(hipStream_t s, const bkh_matrix *A, const CgfGrid &g, const VT *val, const short *cd, T *x,
                      T *r, T *P0, T *P1, T *q, Ctl *c, double *partA, double *partB, int K) {
    const long long n = A->n;
    // kernel B on 16-byte vectors when every array it streams allows (x is the caller's)
    const bool vec = (((uintptr_t)x | (uintptr_t)r | (uintptr_t)P0 | (uintptr_t)P1 | (uintptr_t)q) & 15) == 0;
    for (int j = 0; j < K; j++) {
        if (g.sell) {
            if (g.BSA == 1024) launch_cgf_a_sell<VT, T, 1024>(s, A, g, r, P0, P1, q, c, j, partB, partA);
            else launch_cgf_a_sell<VT, T, 256>(s, A, g, r, P0, P1, q, c, j, partB, partA);
        } else if (cd)
            hipLaunchKernelGGL((k_cgf_a<VT, T, true>), dim3(g.GA), dim3(64 * CGF_WPB), 0, s, A->nwblk, g.chunk,
                               A->rp, A->wblk, A->ci, cd, val, (const T *)r, P0, P1, q, c, j,
                               (const double *)partB, g.GB, partA);
        else
            hipLaunchKernelGGL((k_cgf_a<VT, T, false>), dim3(g.GA), dim3(64 * CGF_WPB), 0, s, A->nwblk, g.chunk,
                               A->rp, A->wblk, A->ci, cd, val, (const T *)r, P0, P1, q, c, j,
                               (const double *)partB, g.GB, partA);
        if (vec)
            hipLaunchKernelGGL((k_cgf_b<T, true>), dim3(g.GB), dim3(512), 0, s, n, x, r, (const T *)P0,
                               (const T *)P1, (const T *)q, c, j, (const double *)partA, g.GA, partB);
        else
            hipLaunchKernelGGL((k_cgf_b<T, false>), dim3(g.GB), dim3(512), 0, s, n, x, r, (const T *)P0,
                               (const T *)P1, (const T *)q, c, j, (const double *)partA, g.GA, partB);
    }
    hipLaunchKernelGGL(k_cgf_tail, dim3(1), dim3(64), 0, s, c, K, (const double *)partB, g.GB);
}

// Graph of K fused iterations plus the tail.
template <typename VT, typename T>
static hipGraphExec_t cgf_capture(ThreadCtx *t, hipStream_t s, const bkh_matrix *A, const VT *val, const short *cd,
                                  T *x, T *r, T *P0, T *P1, T *q, Ctl *c, double *partA, double *partB, int K) {
    const CgfGrid g = cgf_grid(A);
    GraphKey key = {0xC6F0, (uintptr_t)A, (uintptr_t)val, (uintptr_t)cd, (uintptr_t)x, (uintptr_t)r, (uintptr_t)P0,
                    (uintptr_t)P1, (uintptr_t)q, (uintptr_t)c, (uintptr_t)partA, (uintptr_t)partB,
                    (uintptr_t)sizeof(T), (uintptr_t)K, (uintptr_t)(g.BSA * 65536 + g.GA)};
    return capture_cached(t, key, s, 1, [=]() {   // by value: eager mode runs the body after return
        cgf_issue<VT, T>(s, A, g, val, cd, x, r, P0, P1, q, c, partA, partB, K);
    });
}

// Start of the fused path: r.r of iteration 0 in its parity slot, beta_0 = 0.
struct CgfInitFin {
    __device__ void operator()(Ctl *c, const double *tot) const {
        c->s[CGF_RR] = tot[0];
        c->s[CGF_FRESH] = 1.0;
        c->fnb = sqrt(tot[1]);
        if (!(c->fnb > 0.0)) ctl_breakdown(c);
    }
};

// Plain FP64 CG; x updated in place. P holds 2n entries (the two p buffers),
// part the init partials, partA/partB CGF_NPART each.
static int cg_fp64(ThreadCtx *t, const bkh_matrix *A, const double *b, double *x, double *r, double *P,
                   double *q, Ctl *c, double *part, double *partA, double *partB, double tol, int max_it,
                   int K) {
    hipStream_t s = t->stream;
    const long long n = A->n;
    const int G = vec_grid(n, 256);
    const bool fused = cg_can_fuse(A);
    if (ctl_init(t, c, tol, max_it)) return BKH_DEVICE_ERROR;
    Epi<0> e0;
    memset(&e0, 0, sizeof e0);
    launch_spmv<double, double, 0>(s, A, A->val, x, q, 1.0, 0.0, e0, nullptr, nullptr, NoFin());
    hipLaunchKernelGGL(k_cg_init<double>, dim3(G), dim3(256), 0, s, n, b, q, r, P, (int)fused, part);
    if (fused) launch_finalize<2>(s, part, G, c, CgfInitFin());
    else launch_finalize<2>(s, part, G, c, CgInitFin());
    LAUNCH_OK();
    hipGraphExec_t ge;
    if (fused) {
        ge = cgf_capture<double, double>(t, s, A, A->val, nullptr, x, r, P, P + n, q, c, partA, partB, K);
    } else {
        ge = capture_cached(t, {0xC6F1, (uintptr_t)A, (uintptr_t)x, (uintptr_t)r, (uintptr_t)P, (uintptr_t)q,
                                (uintptr_t)c, (uintptr_t)part}, s, K, [&]() {
            launch_spmv_dot_self<double, double>(s, A, A->val, P, q, c, part, CgAlphaFin());
            hipLaunchKernelGGL(k_cg_xr<double>, dim3(G), dim3(256), 0, s, n, x, r, P, q, c, part);
            launch_finalize<1>(s, part, G, c, CgBetaFin());
            hipLaunchKernelGGL(k_cg_p<double>, dim3(G), dim3(256), 0, s, n, P, r, c);
        });
    }
    if (!ge) { bkh_set_error("hipGraph capture of the CG iteration failed"); return BKH_DEVICE_ERROR; }
    int rc = run_graph_until_done(t, ge, c, max_it / K + 2);
    return rc;
}

// Control block of a running fused solve at its first iteration (base = first = 0,
// r.r slots 1, not fresh), for timing kernel A alone.
__global__ void k_cgf_bench_ctl(Ctl *c) {
    if (threadIdx.x == 0) {
        c->base = 0; c->first = 0; c->fnb = 1.0;
        c->s[CGF_RR] = 1.0; c->s[CGF_RR + 1] = 1.0; c->s[CGF_FRESH] = 0.0;
    }
}

// ------------------------------------------------ FP32 with reliable updates

#define RU_DELTA 0.01

// r32 = (float) r64, x32 = 0, and (first time) the p buffer p0 = 0 or = r32.
__global__ __launch_bounds__(256) void k_ru_load(long long n, const double *r64, float *r32, float *x32,
                                                 float *p0, int p_is_r, double *part) {
    double v[1] = {0.0};
    ROWLOOP(i, n) {
        const float ri = (float)r64[i];
        r32[i] = ri;
        x32[i] = 0.0f;
        if (p0) p0[i] = p_is_r ? ri : 0.0f;
        v[0] += (double)ri * (double)ri;
    }
    block_partial<1, 256>(v, part);
}

// x += x32 (FP64), the reliable update's solution step.
__global__ __launch_bounds__(256) void k_ru_addx(long long n, double *x, const float *x32) {
    ROWLOOP(i, n) x[i] += (double)x32[i];
}

// Inner control after a reliable update: r.r of the new r32, a new target, the
// iteration index continues (the fused path's pass starts at it, with beta from
// the new r.r over the last iterated one); ||b|| comes from the FP64 residual
// (co->s[2]).
struct RuResumeFin {
    double target;
    int first;
    const Ctl *co;
    __device__ void operator()(Ctl *c, const double *tot) const {
        c->s[CG_RR] = tot[0];
        if (first) c->s[CG_BETA] = 0.0;
        c->s[CGF_RR + (c->it & 1)] = tot[0];
        c->s[CGF_FRESH] = first ? 1.0 : 0.0;
        c->base = c->it;
        c->first = c->it;
        c->fnb = co->s[2];
        c->tol = target;
        c->done = 0;
        c->status = ST_RUNNING;
        if (!(c->fnb > 0.0) || !(tot[0] == tot[0])) ctl_breakdown(c);
    }
};

__global__ void k_ctl_resume_gate(Ctl *c) {
    // finalize kernels return early on a stopped control block: clear it first
    if (threadIdx.x == 0) c->done = 0;
}

// ---- single-kernel FP32 inner iteration (SELL matrices)

// x += x32 + the updates the stopped kernel k owed x: alpha_{k-1} p_{k-1}, and for an
// odd k also alpha_{k-2} p_{k-2}, which kernel k-1 deferred (p_{k-2} is still in the
// buffer kernel k would have overwritten); x32 = 0.
__global__ __launch_bounds__(256) void k_cg1_fold(long long n, double *x, float *x32, const float *P,
                                                  const Ctl *c) {
    const int k = c->it;
    const bool stopped = c->status != ST_RUNNING;
    const double al = stopped ? c->s[CG1_ALPHA] : 0.0;
    const double ad = (stopped && (k & 1) && k != c->first) ? c->s[CG1_ALPHAD] : 0.0;
    const float *p = P + (size_t)((k + 1) & 1) * n, *p2 = P + (size_t)(k & 1) * n;
    ROWLOOP(i, n) {
        const double d = ad != 0.0 ? ad * (double)p2[i] : 0.0;
        x[i] += (double)x32[i] + d + al * (double)p[i];
        x32[i] = 0.0f;
    }
}

// Pass start: the replaced residual r32 = r64 goes where kernel k reads r_{k-1}
// (alpha = 0 there); the first pass also clears x32 and the q, p buffers it reads.
__global__ __launch_bounds__(256) void k_cg1_load(long long n, const double *r64, float *x32, float *R,
                                                  float *Q, float *P, int first_pass, const Ctl *c,
                                                  double *part) {
    const size_t po = first_pass ? (size_t)n : (size_t)((c->it + 1) & 1) * n;
    double v[1] = {0.0};
    ROWLOOP(i, n) {
        const float ri = (float)r64[i];
        R[po + i] = ri;
        if (first_pass) { x32[i] = 0.0f; Q[po + i] = 0.0f; P[po + i] = 0.0f; }
        v[0] += (double)ri * (double)ri;
    }
    block_partial<1, 256>(v, part);
}

// Inner control at a pass start: rho_k of the replaced residual in slot k&1, the
// previous rho stays in the other slot (beta = rho_k / rho_{k-1}, 0 when fresh).
struct Cg1ResumeFin {
    double target;
    int first;
    const Ctl *co;
    __device__ void operator()(Ctl *c, const double *tot) const {
        if (first) { c->it = 0; c->s[CGF_RR + 1] = tot[0]; }
        const int k = c->it;
        c->s[CGF_RR + (k & 1)] = tot[0];
        c->s[CGF_FRESH] = first ? 1.0 : 0.0;
        c->s[CG1_ALPHA] = 0.0;            // nothing deferred until the pass's first kernel
        c->s[CG1_ALPHAD] = 0.0;
        c->base = k;
        c->first = k;
        c->fnb = co->s[2];
        c->tol = target;
        c->done = 0;
        c->status = ST_RUNNING;
        if (!(c->fnb > 0.0) || !(tot[0] == tot[0])) ctl_breakdown(c);
    }
};

// Launch shape of the single-kernel FP32 iteration. Every workgroup first re-sums
// all partials of the previous kernel (3 per workgroup), so large matrices run
// CG1_NPART workgroups of 512 threads (8 waves, four slices per wave at 128^3):
// measured 15.0 ms per 128^3 solve against 16.2 ms with 2048 workgroups of 256
// threads, whose partials cost ~100 MB of L2 reads per kernel; 1024 workgroups of
// 1024 threads were slower again (17.6 ms).
#define CG1_NPART 1024
struct Cg1Grid { int G, BS, chunk; };

static Cg1Grid cg1_grid(const bkh_matrix *A) {
    Cg1Grid g;
    if (A->nsl >= 8 * CG1_NPART) {
        g.BS = 512;
        g.G = CG1_NPART;
    } else {
        g.BS = 256;
        g.G = std::max(1, std::min((A->nsl + 3) / 4, CGF_NPART));
    }
    g.chunk = (A->nsl + g.G - 1) / g.G;
    return g;
}

template <int BS>
static void launch_cg1(hipStream_t s, const bkh_matrix *A, const Cg1Grid &g, float *x32, float *R, float *Q,
                       float *P, Ctl *c, int j, const double *pin, double *pout) {
    if (cg_dia(A) && A->dval32)
        hipLaunchKernelGGL((k_cg1_sell<float, int, BS, true>), dim3(g.G), dim3(BS), 0, s, (int)A->n, A->nsl, g.chunk,
                           A->rp, A->dso, (const float *)A->dval32, (const int *)A->doff, x32, R, Q, P, c, j,
                           pin, g.G, pout);
    else if (A->scd)
        hipLaunchKernelGGL((k_cg1_sell<float, short, BS>), dim3(g.G), dim3(BS), 0, s, (int)A->n, A->nsl, g.chunk,
                           A->rp, A->sso, (const float *)A->sval32, (const short *)A->scd, x32, R, Q, P, c, j,
                           pin, g.G, pout);
    else
        hipLaunchKernelGGL((k_cg1_sell<float, int, BS>), dim3(g.G), dim3(BS), 0, s, (int)A->n, A->nsl, g.chunk,
                           A->rp, A->sso, (const float *)A->sval32, (const int *)A->scol, x32, R, Q, P, c, j,
                           pin, g.G, pout);
}

static hipGraphExec_t cg1_capture(ThreadCtx *t, hipStream_t s, const bkh_matrix *A, float *x32, float *R,
                                  float *Q, float *P, Ctl *c, double *part2, int K) {
    const Cg1Grid g = cg1_grid(A);
    GraphKey key = {0xC6F3, (uintptr_t)A, (uintptr_t)x32, (uintptr_t)R, (uintptr_t)Q, (uintptr_t)P, (uintptr_t)c,
                    (uintptr_t)part2, (uintptr_t)K, (uintptr_t)(g.BS * 65536 + g.G)};
    return capture_cached(t, key, s, 1, [=]() {   // by value: eager mode runs the body after return
        for (int j = 0; j < K; j++) {
            // kernel base+j reads the partials of base+j-1: parity of j (K even)
            const double *pin = part2 + (size_t)((j + 1) & 1) * 3 * CGF_NPART;
            double *pout = part2 + (size_t)(j & 1) * 3 * CGF_NPART;
            if (g.BS == 512) launch_cg1<512>(s, A, g, x32, R, Q, P, c, j, pin, pout);
            else launch_cg1<256>(s, A, g, x32, R, Q, P, c, j, pin, pout);
        }
        hipLaunchKernelGGL(k_cg1_tail, dim3(1), dim3(256), 0, s, c, K,
                           (const double *)(part2 + (size_t)((K + 1) & 1) * 3 * CGF_NPART), g.G);
    });
}

static int cg_reliable1(ThreadCtx *t, const bkh_matrix *A, const double *b, double *x, double tol,
                        int max_iter, int K, double delta, bkh_info *inf) {
    hipStream_t s = t->stream;
    const long long n = A->n;
    const int G = vec_grid(n, 256);
    if (K & 1) K++;                       // partial buffers alternate by the parity of j
    const int nparts = std::max(spmv_grid(A), G) * 2;
    size_t bytes = 2 * carve_size(sizeof(Ctl)) + carve_size(sizeof(double) * nparts) +
                   carve_size(sizeof(double) * 6 * CGF_NPART) + carve_size(sizeof(double) * n) +
                   7 * carve_size(sizeof(float) * n);
    void *ws = bkh_workspace(bytes);
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *co = cv.take<Ctl>(1), *ci = cv.take<Ctl>(1);
    double *part = cv.take<double>(nparts);
    double *part2 = cv.take<double>(6 * CGF_NPART);
    double *r64 = cv.take<double>(n);
    float *x32 = cv.take<float>(n), *R = cv.take<float>(2 * n), *Q = cv.take<float>(2 * n), *P = cv.take<float>(2 * n);
    if (ctl_init(t, co, 0.0, 1) || ctl_init(t, ci, 1.0, max_iter)) return BKH_DEVICE_ERROR;
    hipGraphExec_t ge = cg1_capture(t, s, A, x32, R, Q, P, ci, part2, K);
    if (!ge) { bkh_set_error("hipGraph capture of the FP32 CG iteration failed"); return BKH_DEVICE_ERROR; }
    int rc = 0, updates = 0;
    int used = 0, inner_status = ST_RUNNING;   // the inner control after the last pass
    double rel = 0.0;
    for (;;) {
        if (updates > 0)
            hipLaunchKernelGGL(k_cg1_fold, dim3(G), dim3(256), 0, s, n, x, x32, (const float *)P, (const Ctl *)ci);
        if ((rc = resid_fp64(t, A, 0.0, b, x, r64, co, part))) break;
        if ((rc = ctl_fetch(t, co))) break;
        rel = t->ctl_host->s[0];
        if (rel < tol || !(rel == rel)) break;
        if (updates > 0 && (inner_status == ST_BREAKDOWN || used >= max_iter)) break;
        const double target = std::max(0.5 * tol, delta * rel);
        hipLaunchKernelGGL(k_cg1_load, dim3(G), dim3(256), 0, s, n, (const double *)r64, x32, R, Q, P,
                           (int)(updates == 0), (const Ctl *)ci, part);
        hipLaunchKernelGGL(k_ctl_resume_gate, dim3(1), dim3(64), 0, s, ci);
        launch_finalize<1>(s, part, G, ci, Cg1ResumeFin{target, updates == 0, co});
        LAUNCH_OK();
        // leaves the inner control block in t->ctl_host (fold and resid do not change it)
        if ((rc = run_graph_until_done(t, ge, ci, (max_iter - used) / K + 2))) break;
        used = t->ctl_host->it;
        inner_status = t->ctl_host->status;
        updates++;
    }
    if (rc) return rc;
    if (ctl_fetch(t, ci)) return BKH_DEVICE_ERROR;
    const int iters = t->ctl_host->it;
    inf->inner_iters = iters;
    inf->passes = updates;
    inf->rec_res = rel;
    if (rel < tol) {
        inf->status = ST_CONVERGED;
        inf->iters = iters > 0 ? iters - 1 : 0;
        return inf->iters;
    }
    inf->status = (rel == rel && t->ctl_host->status != ST_BREAKDOWN) ? ST_MAXIT : ST_BREAKDOWN;
    return inf->status == ST_MAXIT ? -1 : -2;
}

static int cg_reliable(ThreadCtx *t, const bkh_matrix *A, const double *b, double *x, double tol,
                       int max_iter, int K, double delta, bkh_info *inf) {
    hipStream_t s = t->stream;
    const long long n = A->n;
    const int G = vec_grid(n, 256);
    const bool fused = cg_can_fuse(A);
    const int nparts = std::max(spmv_grid(A), G) * 2;
    size_t bytes = 2 * carve_size(sizeof(Ctl)) + carve_size(sizeof(double) * nparts) +
                   2 * carve_size(sizeof(double) * CGF_NPART) + carve_size(sizeof(double) * n) +
                   5 * carve_size(sizeof(float) * n);
    void *ws = bkh_workspace(bytes);
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *co = cv.take<Ctl>(1), *ci = cv.take<Ctl>(1);
    double *part = cv.take<double>(nparts);
    double *partA = cv.take<double>(CGF_NPART), *partB = cv.take<double>(CGF_NPART);
    double *r64 = cv.take<double>(n);
    float *x32 = cv.take<float>(n), *r32 = cv.take<float>(n), *P = cv.take<float>(2 * n), *q = cv.take<float>(n);
    const short *cd = A->cd16;
    if (ctl_init(t, co, 0.0, 1) || ctl_init(t, ci, 1.0, max_iter)) return BKH_DEVICE_ERROR;
    hipGraphExec_t ge;
    if (fused) {
        ge = cgf_capture<float, float>(t, s, A, A->val32, cd, x32, r32, P, P + n, q, ci, partA, partB, K);
    } else {
        ge = capture_cached(t, {0xC6F2, (uintptr_t)A, (uintptr_t)x32, (uintptr_t)r32, (uintptr_t)P, (uintptr_t)q,
                                (uintptr_t)ci, (uintptr_t)part}, s, K, [&]() {
            launch_spmv_dot_self<float, float>(s, A, A->val32, P, q, ci, part, CgAlphaFin());
            hipLaunchKernelGGL(k_cg_xr<float>, dim3(G), dim3(256), 0, s, n, x32, r32, P, q, ci, part);
            launch_finalize<1>(s, part, G, ci, CgBetaFin());
            hipLaunchKernelGGL(k_cg_p<float>, dim3(G), dim3(256), 0, s, n, P, r32, ci);
        });
    }
    if (!ge) { bkh_set_error("hipGraph capture of the FP32 CG iteration failed"); return BKH_DEVICE_ERROR; }
    int rc = 0, updates = 0;
    double rel = 0.0;
    for (;;) {
        // reliable update: x += x32; r = b - A x in FP64; r32 = r, x32 = 0
        if (updates > 0) hipLaunchKernelGGL(k_ru_addx, dim3(G), dim3(256), 0, s, n, x, (const float *)x32);
        if ((rc = resid_fp64(t, A, 0.0, b, x, r64, co, part))) break;
        if ((rc = ctl_fetch(t, co))) break;
        rel = t->ctl_host->s[0];
        if (rel < tol || !(rel == rel)) break;
        if ((rc = ctl_fetch(t, ci))) break;
        const int used = t->ctl_host->it;
        if (updates > 0 && (t->ctl_host->status == ST_BREAKDOWN || used >= max_iter)) break;
        const double target = std::max(0.5 * tol, delta * rel);
        // first pass: the p buffer read by the first iteration is p0 = 0 (fused,
        // beta = 0 gives p = r) or p = r (unfused)
        hipLaunchKernelGGL(k_ru_load, dim3(G), dim3(256), 0, s, n, (const double *)r64, r32, x32,
                           updates == 0 ? P : nullptr, (int)!fused, part);
        hipLaunchKernelGGL(k_ctl_resume_gate, dim3(1), dim3(64), 0, s, ci);
        launch_finalize<1>(s, part, G, ci, RuResumeFin{target, updates == 0, co});
        // unfused path: the stopped iteration skipped its direction update p = r + beta p;
        // apply it with the new r32 (the fused path forms p inside its next SpMV)
        if (!fused && updates > 0)
            hipLaunchKernelGGL(k_cg_p<float>, dim3(G), dim3(256), 0, s, n, P, (const float *)r32, ci);
        LAUNCH_OK();
        if ((rc = run_graph_until_done(t, ge, ci, (max_iter - used) / K + 2))) break;
        updates++;
    }
    if (rc) return rc;
    if (ctl_fetch(t, ci)) return BKH_DEVICE_ERROR;
    const int iters = t->ctl_host->it;
    inf->inner_iters = iters;
    inf->passes = updates;
    inf->rec_res = rel;
    if (rel < tol) {
        inf->status = ST_CONVERGED;
        inf->iters = iters > 0 ? iters - 1 : 0;
        return inf->iters;
    }
    inf->status = (rel == rel && t->ctl_host->status != ST_BREAKDOWN) ? ST_MAXIT : ST_BREAKDOWN;
    return inf->status == ST_MAXIT ? -1 : -2;
}

extern "C" int bkh_cg(bkh_matrix *A, const double *b, double *x, double tol, int max_iter,
                      const bkh_opts *opts, bkh_info *info) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !b || !x) return BKH_BAD_ARGUMENT;
    bkh_opts o = opts ? *opts : bkh_opts{};
    bkh_info inf;
    memset(&inf, 0, sizeof inf);
    const long long n = A->n;
    int K = o.graph_iters > 0 ? o.graph_iters : default_graph_iters(n);
    if (max_iter <= 0) {                     // cg.c:15,67-70: error = 0 -> returns 0
        inf.ret = 0; inf.status = ST_CONVERGED;
        if (info) *info = inf;
        return 0;
    }
    int ret;
    HIP_OK(hipEventRecord(t->ev0, t->stream));
    if (o.precision == BKH_FP32_IR) {
        if (!A->val32) { bkh_set_error("FP32_IR needs a matrix created with want_fp32"); return BKH_BAD_ARGUMENT; }
        const double delta = o.inner_tol > 0 ? o.inner_tol : RU_DELTA;
        ret = (cgf_sell(A) && A->sval32 && o.path != 1) ? cg_reliable1(t, A, b, x, tol, max_iter, K, delta, &inf)
                                                           : cg_reliable(t, A, b, x, tol, max_iter, K, delta, &inf);
        if (ret < -2) return ret;
    } else {
        const int nparts = std::max(spmv_grid(A), vec_grid(n, 256)) * 2;
        size_t bytes = carve_size(sizeof(Ctl)) + carve_size(sizeof(double) * nparts) +
                       2 * carve_size(sizeof(double) * CGF_NPART) + 4 * carve_size(sizeof(double) * n);
        void *ws = bkh_workspace(bytes);
        if (!ws) return BKH_DEVICE_ERROR;
        Carve cv(ws);
        Ctl *c = cv.take<Ctl>(1);
        double *part = cv.take<double>(nparts);
        double *partA = cv.take<double>(CGF_NPART), *partB = cv.take<double>(CGF_NPART);
        double *r = cv.take<double>(n), *P = cv.take<double>(2 * n), *q = cv.take<double>(n);
        int rc = cg_fp64(t, A, b, x, r, P, q, c, part, partA, partB, tol, max_iter, K);
        if (rc) return rc;
        ret = ctl_result(t->ctl_host);
        inf.status = t->ctl_host->status;
        inf.iters = t->ctl_host->result_it;
        inf.rec_res = t->ctl_host->err;
    }
    HIP_OK(hipEventRecord(t->ev1, t->stream));
    HIP_OK(hipEventSynchronize(t->ev1));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    inf.time_ms = ms;
    bkh_note_solve_ms(ms);
    inf.ret = ret;
    if (info) {
        double tr = 0.0;
        if (true_residual(t, A, 0.0, b, x, &tr) == 0) inf.true_res = tr;
        *info = inf;
    }
    return ret;
}

// Average ms of the fused CG kernel A (SpMV with the direction update in its gather,
// the dominant kernel of a CG iteration) over reps launches, HIP events on the
// launching stream. Vectors are zero; the control block holds a running solve.
extern "C" int bkh_bench_cg_kernel(bkh_matrix *A, int fp32, int reps, double *avg_ms) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || reps < 1 || !cgf_sell(A) || (fp32 && !A->sval32)) {
        bkh_set_error("the fused CG kernel needs a matrix on the SELL layout");
        return BKH_BAD_ARGUMENT;
    }
    const long long n = A->n;
    const size_t vb = fp32 ? sizeof(float) : sizeof(double);
    void *ws = bkh_workspace(carve_size(sizeof(Ctl)) + 2 * carve_size(sizeof(double) * CGF_NPART) +
                             4 * carve_size(vb * n));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1);
    double *partA = cv.take<double>(CGF_NPART), *partB = cv.take<double>(CGF_NPART);
    char *r = cv.take<char>(vb * n), *P = cv.take<char>(2 * vb * n), *q = cv.take<char>(vb * n);
    hipStream_t s = t->stream;
    HIP_OK(hipMemsetAsync(r, 0, vb * n, s));
    HIP_OK(hipMemsetAsync(P, 0, 2 * vb * n, s));
    if (ctl_init(t, c, 0.0, 1 << 30)) return BKH_DEVICE_ERROR;
    hipLaunchKernelGGL(k_cgf_bench_ctl, dim3(1), dim3(64), 0, s, c);
    const CgfGrid g = cgf_grid(A);
    auto launch = [&]() {
        if (fp32) {
            float *r_ = (float *)r, *p0 = (float *)P, *p1 = (float *)P + n, *q_ = (float *)q;
            if (g.BSA == 1024) launch_cgf_a_sell<float, float, 1024>(s, A, g, r_, p0, p1, q_, c, 0, partB, partA);
            else launch_cgf_a_sell<float, float, 256>(s, A, g, r_, p0, p1, q_, c, 0, partB, partA);
        } else {
            double *r_ = (double *)r, *p0 = (double *)P, *p1 = (double *)P + n, *q_ = (double *)q;
            if (g.BSA == 1024) launch_cgf_a_sell<double, double, 1024>(s, A, g, r_, p0, p1, q_, c, 0, partB, partA);
            else launch_cgf_a_sell<double, double, 256>(s, A, g, r_, p0, p1, q_, c, 0, partB, partA);
        }
    };
    launch();
    HIP_OK(hipEventRecord(t->ev0, s));
    for (int k = 0; k < reps; k++) launch();
    HIP_OK(hipEventRecord(t->ev1, s));
    LAUNCH_OK();
    HIP_OK(hipEventSynchronize(t->ev1));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    *avg_ms = ms / reps;
    return 0;
}

// ------------------------------------------------------------ allocation-free CG
//
// bkh_cg_noalloc: plain FP64 CG (the fused two-kernel iteration above) whose whole
// workspace -- control blocks, partial sums, r, the two p buffers, q -- lives in the
// file-scope device array `scratch` below, so the call allocates nothing: no
// hipMalloc/hipFree, no graph instantiation, no stream or event creation. The kernels
// are issued K iterations at a time on the null stream and the host reads the
// control block between groups (hipMemcpy into a stack copy). Every call on one GPU
// uses that GPU's single copy of the array: calls that overlap on one GPU race, and
// each call initialises everything it reads. Rows are limited to CGNA_MAX_ROWS.

#define CGNA_MAX_ROWS (1 << 20)
#define CGNA_MAX_PARTS 16384            // doubles for the initial reductions' partials
#define CGNA_MAX_DEV 64
static constexpr size_t CGNA_BYTES =
    2 * (((sizeof(Ctl) + 255) & ~(size_t)255) + 256) + (sizeof(double) * CGNA_MAX_PARTS + 256) +
    2 * (sizeof(double) * CGF_NPART + 256) + 4 * (sizeof(double) * (size_t)CGNA_MAX_ROWS + 256) + 256;
static __device__ __align__(256) double scratch[CGNA_BYTES / sizeof(double)];

__global__ void k_ctl_init(Ctl *c, double tol, int max_it, int it0);
__global__ __launch_bounds__(256) void k_resid(long long n, const double *b, const double *x, double sigma,
                                               double *r, Ctl *c, double *part);

struct CgnaResidFin {   // ||b - A x|| / ||b|| into s[0]
    __device__ void operator()(Ctl *c, const double *tot) const { c->s[0] = sqrt(tot[0]) / sqrt(tot[1]); }
};

// The device address of `scratch` on the current device, looked up once per device
// (hipGetSymbolAddress with that device current); nullptr with the error set.
static double *cgna_scratch() {
    static std::mutex mu;
    static double *addr[CGNA_MAX_DEV];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= CGNA_MAX_DEV) {
        bkh_set_error("allocation-free CG: no current HIP device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            bkh_set_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
            return nullptr;
        }
        addr[dev] = (double *)p;
    }
    return addr[dev];
}

extern "C" long long bkh_cg_noalloc_max_rows(void) { return CGNA_MAX_ROWS; }

extern "C" int bkh_cg_noalloc(const bkh_matrix *A, const double *b, double *x, double tol, int max_iter,
                              bkh_info *info) {
    if (!A || !b || !x) return BKH_BAD_ARGUMENT;
    const long long n = A->n;
    if (n > CGNA_MAX_ROWS) {
        bkh_set_error("allocation-free CG: %lld rows exceed the scratch limit of %d", n, CGNA_MAX_ROWS);
        return BKH_BAD_ARGUMENT;
    }
    if (!cg_can_fuse(A)) {
        bkh_set_error("allocation-free CG: needs the fused iteration (SELL-64 layout or rows of <= %d nonzeros)",
                      SPMV_WCAP);
        return BKH_BAD_ARGUMENT;
    }
    bkh_info inf;
    memset(&inf, 0, sizeof inf);
    if (max_iter <= 0) {                     // cg.c:15,67-70: error = 0 -> returns 0
        inf.status = ST_CONVERGED;
        if (info) *info = inf;
        return 0;
    }
    const int G = vec_grid(n, 256);
    const int nparts = std::max(spmv_grid(A), G) * 2;
    if (nparts > CGNA_MAX_PARTS) {
        bkh_set_error("allocation-free CG: %d partial sums exceed the scratch limit", nparts);
        return BKH_BAD_ARGUMENT;
    }
    double *base = cgna_scratch();
    if (!base) return BKH_DEVICE_ERROR;
    Carve cv(base);
    Ctl *c = cv.take<Ctl>(1), *cr = cv.take<Ctl>(1);
    double *part = cv.take<double>(CGNA_MAX_PARTS);
    double *partA = cv.take<double>(CGF_NPART), *partB = cv.take<double>(CGF_NPART);
    double *r = cv.take<double>(n), *P = cv.take<double>(2 * n), *q = cv.take<double>(n);
    if (cv.off > CGNA_BYTES) {
        bkh_set_error("allocation-free CG: scratch layout overflow");
        return BKH_DEVICE_ERROR;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const hipStream_t s = 0;
    const int K = default_graph_iters(n);
    const CgfGrid g = cgf_grid(A);
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(64), 0, s, c, tol, max_iter, 0);
    Epi<0> e0;
    memset(&e0, 0, sizeof e0);
    launch_spmv<double, double, 0>(s, A, A->val, x, q, 1.0, 0.0, e0, nullptr, nullptr, NoFin());
    hipLaunchKernelGGL(k_cg_init<double>, dim3(G), dim3(256), 0, s, n, b, (const double *)q, r, P, 1, part);
    launch_finalize<2>(s, part, G, c, CgfInitFin());
    LAUNCH_OK();
    Ctl h;
    const int max_groups = max_iter / K + 2;
    for (int grp = 0; grp < max_groups; grp++) {
        cgf_issue<double, double>(s, A, g, A->val, nullptr, x, r, P, P + n, q, c, partA, partB, K);
        LAUNCH_OK();
        HIP_OK(hipMemcpy(&h, c, sizeof(Ctl), hipMemcpyDeviceToHost));
        if (h.done) break;
    }
    // FP64 true residual into q (the iteration is over)
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(64), 0, s, cr, 0.0, 1, 0);
    launch_spmv<double, double, 0>(s, A, A->val, x, q, 1.0, 0.0, e0, nullptr, nullptr, NoFin());
    hipLaunchKernelGGL(k_resid, dim3(G), dim3(256), 0, s, n, b, (const double *)x, 0.0, q, cr, part);
    launch_finalize<2>(s, part, G, cr, CgnaResidFin());
    LAUNCH_OK();
    Ctl hr;
    HIP_OK(hipMemcpy(&hr, cr, sizeof(Ctl), hipMemcpyDeviceToHost));
    inf.ret = ctl_result(&h);
    inf.status = h.done ? h.status : ST_MAXIT;
    inf.iters = h.result_it;
    inf.rec_res = h.err;
    inf.true_res = hr.s[0];
    inf.time_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (info) *info = inf;
    return inf.ret;
}
