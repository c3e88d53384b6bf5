// batched.hip -- many independent systems in one launch per phase (or one persistent
// launch), with an autotuner.
//
// The reference solves one system per call (src/bicgstabl.c:13-138); a batch of
// uneven systems would be that many sequential solves, each a long chain of small
// launches. Here the systems are stacked block-diagonally in one CSR (global row
// offsets, column indices shifted by the system's offset) and every phase of
// BiCGSTAB(l) is ONE launch over the whole batch:
//   * SpMV units: groups of four wave blocks (<= rows_cap rows, <= nnz_cap nonzeros
//     each) that never straddle two systems (each system's list is padded with empty
//     blocks), so a unit's epilogue partial sums belong to one system;
//   * vector tiles: row ranges of <= tile_rows rows inside one system;
//   * per-system finalize: one workgroup per system sums that system's partials in
//     a fixed order and runs its own scalar recurrence (alpha, beta, the MR step,
//     the convergence test) in its BtSys record;
//   * every unit and tile reads its system's `done` flag at entry, so converged
//     systems cost nothing while the rest continue; a one-workgroup kernel at the end
//     of each outer iteration stops the graph when no system is active.
// With BKH_FP32_IR the iteration runs in FP32 on a correction, inside per-system
// FP64 iterative refinement (true residual b - A x in FP64 for every system at once).
// bkh_batch_tune() times candidate SpMV blockings (rows per wavefront, nonzeros per
// wavefront from the batch's row-length histogram), the column-chunked SpMV and vector
// tile sizes on the batch itself and keeps the fastest. With opts path 2 the whole solve
// is one persistent cooperative launch instead (k_bt_persist_lc with the chunked SpMV,
// k_bt_persist with the CSR-stream one), with the same arithmetic.
#include <math.h>
#include <climits>
#include <algorithm>
#include <array>
#include <cmath>
#include <vector>
#include "bkh_host.hpp"
#include "bkh_mr.hpp"
#include "bkh_solver_common.hpp"
#include "bkh_spmv.hpp"

struct BtSys {
    double g0, alpha, beta, oml, fnb, err, tol, rel;
    double cx[MR_D], cr[MR_D], cp[MR_D];
    int done, status, it, result, max_it, pad[3];
};

struct bkh_batch {
    int ns = 0;
    long long N = 0, nnz = 0;
    std::vector<long long> off;     // row offsets, ns + 1
    std::vector<int> rp_h;          // global row pointers (host), N + 1
    int *rp = nullptr, *ci = nullptr;
    double *val = nullptr;
    float *val32 = nullptr;
    int max_row = 0;
    // SpMV units
    int *wblk = nullptr, *usys = nullptr, *ub = nullptr;
    int nunit = 0;
    // vector tiles
    int *trow = nullptr, *tsys = nullptr, *tb = nullptr;
    int ntile = 0;
    int rows_cap = SPMV_WROWS, nnz_cap = SPMV_WCAP, tile_rows = 1024;
    // column-chunked copy (built when every row's columns ascend): items of <= LC_ROWS
    // rows of one system, entries chunk-major, R + 1 entry pointers per chunk
    int lc_ok = 0, use_lc = 0, nitem = 0;
    struct LcItem *items = nullptr;
    double *lval = nullptr;
    float *lval32 = nullptr;
    unsigned short *lcol = nullptr;
    unsigned short *lrp = nullptr;
    int *lcb = nullptr, *ib = nullptr;
};

// Column-chunked batched SpMV. An item is LC_ROWS (or fewer) consecutive rows of one
// system; its entries are stored chunk by chunk (chunk c: the rows' entries whose
// system-local column lies in [c LC_CH, (c+1) LC_CH), row by row in stored order, the
// column as a 16-bit offset into the chunk), with the chunk's first entry (a multiple
// of 8) and R + 1 16-bit entry offsets from it per chunk (an item has as many rows, at
// most LC_ROWS, as keep every chunk's entries below 65536). For
// each chunk the workgroup stages that part of the system's x in LDS, then takes the
// chunk's entries in blocks of BLK: each lane forms BLK / LC_NT consecutive products
// v x into LDS, a lane per row continues its running sum over its segment of the block. The random-column gathers
// become LDS reads; columns ascend within a row, so each row still sums its entries in
// stored order and y is bitwise the CSR-stream kernel's.
#define LC_CH 8192
#define LC_NT 1024
#define LC_RPT 4
#define LC_ROWS (LC_NT * LC_RPT)
struct LcItem {
    int sys, xoff, r0, R, nc, rpb, ns, cbi;   // cbi: index of chunk 0's entry base
};

// ------------------------------------------------------------------ kernels

// Every phase is a device function of the unit / tile / system it serves, so the
// same code runs as one launch per phase (graph path) and inside the persistent
// kernel below. `sh` is workgroup scratch for the partial sums (>= 4 NV doubles).
// threadIdx.x behind an empty asm: expressions of it stay inside the phase that forms
// them (the persistent kernels inline every phase into one loop; hoisted to its entry,
// each phase's per-lane addresses would all stay live and spill).
__device__ __forceinline__ int phase_tid() {
    int t = (int)threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// Sums NV values over each 256-thread group of a workgroup of NGRP such groups (all
// groups call it together); the totals are valid in each group's thread 0 and are
// formed in block_sum<NV, 256>'s order (wave sums, then waves 0..3 in turn). `sh`
// holds NGRP * NV * 4 doubles.
template <int NV, int NGRP>
__device__ __forceinline__ void group_sum(double (&v)[NV], double *sh) {
    if constexpr (NGRP == 1) {
        block_sum<NV, 256>(v, sh);
    } else {
        const int tid = phase_tid(), lane = tid & 63, w = tid >> 6;
        double *gs = sh + (w >> 2) * NV * 4;
#pragma unroll
        for (int i = 0; i < NV; i++) v[i] = wave_sum(v[i]);
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i < NV; i++) gs[i * 4 + (w & 3)] = v[i];
        }
        __syncthreads();
        if ((tid & 255) == 0) {
#pragma unroll
            for (int i = 0; i < NV; i++) {
                double t = gs[i * 4];
                for (int k = 1; k < 4; k++) t += gs[i * 4 + k];
                v[i] = t;
            }
        }
        __syncthreads();
    }
}

template <int NV, int NGRP = 1>
__device__ __forceinline__ void unit_partial(double (&v)[NV], double *part, int u, double *sh, bool valid = true) {
    group_sum<NV, NGRP>(v, sh);
    if ((phase_tid() & 255) == 0 && valid) {
#pragma unroll
        for (int i = 0; i < NV; i++) part[(size_t)u * NV + i] = v[i];
    }
}

// y = A x over unit u's four wave blocks; epilogue dot products of y with the vectors
// in e (one partial row per unit). prod: the workgroup's 4 x SPMV_WCAP products.
template <typename VT, typename T, int NV>
__device__ __forceinline__ void bt_spmv_unit(int u, const int *__restrict__ rp, const int *__restrict__ ci,
                                             const VT *__restrict__ val, const int *__restrict__ wblk,
                                             const int *__restrict__ usys, const BtSys *sy, const T *x, T *y,
                                             const Epi<NV, T> &e, double *part, T (*prod_all)[SPMV_WCAP],
                                             double *sh) {
    if (sy && sy[usys[u]].done) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T *prod = prod_all[w];
    const int b = u * 4 + w;
    double dv[NV > 0 ? NV : 1];
#pragma unroll
    for (int i = 0; i < NV; i++) dv[i] = 0.0;
    const int r0 = wblk[b], r1 = wblk[b + 1];
    if (r1 > r0) {
        const int k0 = rp[r0], k1 = rp[r1];
        constexpr int VEC = 16 / sizeof(VT);
        constexpr int STEPS = SPMV_WCAP / (VEC * 64) + 1;
        typedef VT vvec __attribute__((ext_vector_type(VEC)));
        typedef int ivec __attribute__((ext_vector_type(VEC)));
        const int kb = k0 & ~(VEC - 1);
        vvec vv[STEPS];
        ivec cc[STEPS];
#pragma unroll
        for (int j = 0; j < STEPS; j++) {
            const int k = kb + VEC * (lane + j * 64);
            if (k < k1) {
                vv[j] = *(const vvec *)(val + k);
                cc[j] = *(const ivec *)(ci + k);
            }
        }
#pragma unroll
        for (int j = 0; j < STEPS; j++) {
            const int k = kb + VEC * (lane + j * 64);
#pragma unroll
            for (int q = 0; q < VEC; q++)
                if (k + q >= k0 && k + q < k1) prod[k + q - k0] = mul_rn((T)vv[j][q], x[cc[j][q]]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int row = r0 + lane;
        if (row < r1) {
            const int s0 = rp[row] - k0, s1 = rp[row + 1] - k0;
            T s = (T)0;
            for (int k = s0; k < s1; k++) s = add_rn(s, prod[k]);
            y[row] = s;
#pragma unroll
            for (int i = 0; i < NV; i++)
                dv[i] = epi_operand(e, e.a[i], row, (double)s) * epi_operand(e, e.b[i], row, (double)s);
        }
        // the next unit of this wave rewrites prod only after every lane has read it
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if constexpr (NV > 0) unit_partial<NV>(dv, part, u, sh);
}

template <typename VT, typename T, int NV>
__global__ __launch_bounds__(256) void k_bt_spmv(const int *__restrict__ rp, const int *__restrict__ ci,
                                                 const VT *__restrict__ val, const int *__restrict__ wblk,
                                                 const int *__restrict__ usys, const BtSys *sy, const T *x, T *y,
                                                 Epi<NV, T> e, double *part) {
    __shared__ T prod_all[4][SPMV_WCAP];
    __shared__ double sh[NV * 4 > 0 ? NV * 4 : 1];
    bt_spmv_unit<VT, T, NV>(xcd_block(blockIdx.x, gridDim.x), rp, ci, val, wblk, usys, sy, x, y, e, part, prod_all,
                            sh);
}

template <typename E, int K> struct LcVec { typedef E type __attribute__((ext_vector_type(K))); };

// A lane takes BLK / LC_NT consecutive entries of the block (every chunk's entries start
// at a multiple of 8) with one vector load of values and one of columns, and stores its
// products with one LDS store; the chunk of x is loaded with all of a lane's loads in
// flight before the LDS stores.
template <typename VT>
struct LcArgs {
    const LcItem *items;
    const VT *ev;
    const unsigned short *ec, *erp;
    const int *ecb;
};

// Products per block of the chunked SpMV: the FP64 workgroup's LDS (x chunk + products)
// stays at 80 KB, so two workgroups share a CU.
template <typename T> constexpr int lc_blk() { return sizeof(T) == sizeof(double) ? 2048 : 8192; }

// Item `it` of the chunked SpMV by the whole (LC_NT-thread) workgroup; xs (LC_CH) and
// prod (BLK, 32-byte aligned) are its LDS.
template <typename VT, typename T, int NV, int BLK>
__device__ __forceinline__ void bt_spmv_item(int it, const LcArgs<VT> &la, const BtSys *sy, const T *x, T *y,
                                             const Epi<NV, T> &e, double *part, T *xs, T *prod) {
    constexpr int EPL = BLK / LC_NT, XPT = LC_CH / LC_NT;
    static_assert(EPL >= 2 && EPL <= 8 && BLK % LC_NT == 0, "entries per lane");
    typedef typename LcVec<VT, EPL>::type vvec;
    typedef typename LcVec<unsigned short, EPL>::type cvec;
    typedef typename LcVec<T, EPL>::type pvec;
    static_assert(NV * (LC_NT / 64) * sizeof(double) <= BLK * sizeof(T), "partials fit in prod");
    const VT *__restrict__ ev = la.ev;
    const unsigned short *__restrict__ ec = la.ec;
    const unsigned short *__restrict__ erp = la.erp;
    const int *__restrict__ ecb = la.ecb;
    const LcItem d = la.items[it];
    if (sy && sy[d.sys].done) return;
    const int t = phase_tid();
    T acc[LC_RPT];
#pragma unroll
    for (int q = 0; q < LC_RPT; q++) acc[q] = (T)0;
    for (int c = 0; c < d.nc; c++) {
        const int base = c * LC_CH, len = min(LC_CH, d.ns - base);
        const T *xc = x + d.xoff + base;
        T xr[XPT];
#pragma unroll
        for (int j = 0; j < XPT; j++) {
            const int i = t + LC_NT * j;
            xr[j] = i < len ? xc[i] : (T)0;
        }
        const unsigned short *rp = erp + d.rpb + c * (d.R + 1);
        const int E0 = ecb[d.cbi + c], E1 = E0 + rp[d.R];
        int a[LC_RPT], b[LC_RPT];
        // loads from clamped rows, all in flight before the selects (R >= 1)
        unsigned short ra[LC_RPT], rb[LC_RPT];
#pragma unroll
        for (int q = 0; q < LC_RPT; q++) {
            const int r = min(t + LC_NT * q, d.R - 1);
            ra[q] = rp[r];
            rb[q] = rp[r + 1];
        }
#pragma unroll
        for (int q = 0; q < LC_RPT; q++) {
            const bool in = t + LC_NT * q < d.R;
            a[q] = in ? E0 + ra[q] : 0;
            b[q] = in ? E0 + rb[q] : 0;
            asm volatile("" ::"v"(a[q]), "v"(b[q]));   // keeps these loads ahead of the barrier
        }
        __syncthreads();                                   // the previous chunk's readers are done
#pragma unroll
        for (int j = 0; j < XPT; j++) xs[t + LC_NT * j] = xr[j];
        __syncthreads();
        for (int B0 = E0; B0 < E1; B0 += BLK) {
            const int k = B0 + EPL * t;
            if (k < E1) {
                const vvec v = *(const vvec *)(ev + k);
                const cvec cc = *(const cvec *)(ec + k);
                pvec pr;
#pragma unroll
                for (int i = 0; i < EPL; i++) pr[i] = mul_rn((T)v[i], xs[cc[i]]);
                *(pvec *)(prod + EPL * t) = pr;
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < LC_RPT; q++) {
                const int lo = max(a[q], B0), hi = min(b[q], B0 + BLK);
                for (int j = lo; j < hi; j++) acc[q] = add_rn(acc[q], prod[j - B0]);
            }
            __syncthreads();
        }
    }
    double dv[NV > 0 ? NV : 1];
#pragma unroll
    for (int i = 0; i < NV; i++) dv[i] = 0.0;
#pragma unroll
    for (int q = 0; q < LC_RPT; q++) {
        const int r = t + LC_NT * q;
        if (r < d.R) {
            const int row = d.r0 + r;
            y[row] = acc[q];
#pragma unroll
            for (int i = 0; i < NV; i++)
                dv[i] += epi_operand(e, e.a[i], row, (double)acc[q]) * epi_operand(e, e.b[i], row, (double)acc[q]);
        }
    }
    if constexpr (NV > 0) {
        // prod is free: every lane has passed the barrier after its last read of it
        block_sum<NV, LC_NT>(dv, (double *)prod);
        if (t == 0) {
#pragma unroll
            for (int i = 0; i < NV; i++) part[(size_t)it * NV + i] = dv[i];
        }
    }
}

template <typename VT, typename T, int NV, int BLK>
__global__ __launch_bounds__(LC_NT) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_bt_spmv_lc(LcArgs<VT> la, const BtSys *sy, const T *x, T *y,
                                                      Epi<NV, T> e, double *part) {
    __shared__ T xs[LC_CH];
    __shared__ __attribute__((aligned(32))) T prod[BLK];
    bt_spmv_item<VT, T, NV, BLK>(xcd_block(blockIdx.x, gridDim.x), la, sy, x, y, e, part, xs, prod);
}

// A tile is taken by 256 threads: the whole workgroup of the per-phase kernels, or one
// of the four 256-thread groups of the persistent chunked kernel's workgroups.
#define TILE_ROWS(i, t_)                                                              \
    const int sys = tsys[t_];                                                         \
    for (long long i = (long long)trow[t_] + (phase_tid() & 255); i < trow[t_ + 1]; i += 256)

// r0 = b - A x (A x in r0 on entry); p0 = rs = r0; partials {r0.r0, b.b}.
template <typename T, int NGRP = 1>
__device__ __forceinline__ void bt_init_tile(int t, const int *trow, const int *tsys, const T *b, T *R, T *P,
                                             T *rs, const BtSys *sy, double *part, double *sh, bool valid = true) {
    double v[2] = {0.0, 0.0};
    if (valid && !sy[tsys[t]].done) {
        TILE_ROWS(i, t) {
            const T ri = b[i] - R[i];
            R[i] = ri; P[i] = ri; rs[i] = ri;
            v[0] += (double)ri * (double)ri;
            v[1] += (double)b[i] * (double)b[i];
        }
        (void)sys;
    }
    unit_partial<2, NGRP>(v, part, t, sh, valid);
}
template <typename T>
__global__ __launch_bounds__(256) void k_bt_init(const int *trow, const int *tsys, long long N, const T *b,
                                                 T *R, T *P, T *rs, const BtSys *sy, double *part) {
    __shared__ double sh[8];
    bt_init_tile<T>(blockIdx.x, trow, tsys, b, R, P, rs, sy, part, sh);
}

// The vector tiles step a lane through its rows four at a time (rows i, i + 256,
// i + 512, i + 768): every load of the four is issued before the arithmetic, which then
// runs row by row in the order of the plain loop, so results and partial sums are those
// of the one-row-at-a-time loop.
// U rows per step (template parameter of the tile functions: 4 in the per-phase kernels,
// 1 in the persistent kernels, where the registers of four rows would spill).
#define TILE_ROWS4(t_)                                                                            \
    const long long e_ = trow[(t_) + 1];                                                          \
    for (long long i0_ = (long long)trow[t_] + (phase_tid() & 255); i0_ < e_; i0_ += 256 * U)
// the rows of one step and whether each is in the tile
template <int U>
__device__ __forceinline__ void tile_step(long long i0, long long e, long long (&ii)[U], bool (&in)[U]) {
#pragma unroll
    for (int u = 0; u < U; u++) {
        ii[u] = i0 + 256 * u;
        in[u] = ii[u] < e;
    }
}

// p_q = beta p_q + r_q, q <= j (bicgstabl.c:64-67)
template <typename T, int U = 4>
__device__ __forceinline__ void bt_p_tile(int t, const int *trow, const int *tsys, long long N, int j, T *P,
                                          const T *R, const BtSys *sy) {
    const int sys = tsys[t];
    if (sy[sys].done) return;
    const T be = (T)sy[sys].beta;
    TILE_ROWS4(t) {
        long long ii[U];
        bool in[U];
        tile_step<U>(i0_, e_, ii, in);
        for (int q = 0; q <= j; q++) {
            T pv[U] = {}, rv[U] = {};
#pragma unroll
            for (int u = 0; u < U; u++)
                if (in[u]) { pv[u] = P[q * N + ii[u]]; rv[u] = R[q * N + ii[u]]; }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (in[u]) P[q * N + ii[u]] = add_rn(mul_rn(be, pv[u]), rv[u]);
        }
    }
}
template <typename T>
__global__ __launch_bounds__(256) void k_bt_p(const int *trow, const int *tsys, long long N, int j, T *P,
                                              const T *R, const BtSys *sy) {
    bt_p_tile<T>(blockIdx.x, trow, tsys, N, j, P, R, sy);
}

// r_q -= alpha p_{q+1}, q <= j; x += alpha p_0 (:73-75, :79)
template <typename T, int U = 4>
__device__ __forceinline__ void bt_r_tile(int t, const int *trow, const int *tsys, long long N, int j, const T *P,
                                          T *R, T *x, const BtSys *sy) {
    const int sys = tsys[t];
    if (sy[sys].done) return;
    const T al = (T)sy[sys].alpha;
    TILE_ROWS4(t) {
        long long ii[U];
        bool in[U];
        tile_step<U>(i0_, e_, ii, in);
        for (int q = 0; q <= j; q++) {
            T pv[U] = {}, rv[U] = {};
#pragma unroll
            for (int u = 0; u < U; u++)
                if (in[u]) { pv[u] = P[(q + 1) * N + ii[u]]; rv[u] = R[q * N + ii[u]]; }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (in[u]) R[q * N + ii[u]] = fma(-al, pv[u], rv[u]);
        }
        T pv[U] = {}, xv[U] = {};
#pragma unroll
        for (int u = 0; u < U; u++)
            if (in[u]) { pv[u] = P[ii[u]]; xv[u] = x[ii[u]]; }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (in[u]) x[ii[u]] = fma(al, pv[u], xv[u]);
    }
}
template <typename T>
__global__ __launch_bounds__(256) void k_bt_r(const int *trow, const int *tsys, long long N, int j, const T *P,
                                              T *R, T *x, const BtSys *sy) {
    bt_r_tile<T>(blockIdx.x, trow, tsys, N, j, P, R, x, sy);
}

// MR update of x, r_0, p_0 with partials {r0.r0, rs.r0} (:113-125)
template <typename T, int NGRP = 1, int U = 4>
__device__ __forceinline__ void bt_upd_tile(int t, const int *trow, const int *tsys, long long N, int l, T *P,
                                            T *R, T *x, const T *rs, const BtSys *sy, double *part, double *sh,
                                            bool valid = true) {
    double v[2] = {0.0, 0.0};
    const int sy_i = valid ? tsys[t] : 0;
    if (valid && !sy[sy_i].done) {
        double cx[MR_D], cr[MR_D], cp[MR_D];
        for (int k = 0; k <= l; k++) { cx[k] = sy[sy_i].cx[k]; cr[k] = sy[sy_i].cr[k]; cp[k] = sy[sy_i].cp[k]; }
        TILE_ROWS4(t) {
            long long ii[U];
            bool in[U];
            tile_step<U>(i0_, e_, ii, in);
            double xi[U] = {}, r0[U] = {}, p0[U] = {}, rsv[U] = {};
#pragma unroll
            for (int u = 0; u < U; u++)
                if (in[u]) {
                    xi[u] = (double)x[ii[u]]; r0[u] = (double)R[ii[u]]; p0[u] = (double)P[ii[u]];
                    rsv[u] = (double)rs[ii[u]];
                }
#pragma unroll
            for (int u = 0; u < U; u++) xi[u] += cx[0] * r0[u];
            for (int k = 1; k <= l; k++) {
                double rk[U] = {}, pk[U] = {};
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (in[u]) { rk[u] = (double)R[k * N + ii[u]]; pk[u] = (double)P[k * N + ii[u]]; }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    if (k < l) xi[u] += cx[k] * rk[u];
                    r0[u] -= cr[k] * rk[u];
                    p0[u] -= cp[k] * pk[u];
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (in[u]) {
                    x[ii[u]] = (T)xi[u]; R[ii[u]] = (T)r0[u]; P[ii[u]] = (T)p0[u];
                    const double r0t = (double)(T)r0[u];
                    v[0] += r0t * r0t;
                    v[1] += rsv[u] * r0t;
                }
        }
    }
    unit_partial<2, NGRP>(v, part, t, sh, valid);
}
template <typename T>
__global__ __launch_bounds__(256) void k_bt_upd(const int *trow, const int *tsys, long long N, int l, T *P, T *R,
                                                T *x, const T *rs, const BtSys *sy, double *part) {
    __shared__ double sh[8];
    bt_upd_tile<T>(blockIdx.x, trow, tsys, N, l, P, R, x, rs, sy, part, sh);
}

enum { BT_INIT = 0, BT_ALPHA = 1, BT_GAMMA = 2, BT_MR = 3, BT_UPD = 4, BT_RESID = 5 };

__device__ __forceinline__ void bt_stop(BtSys &s, int status) {
    s.done = 1; s.status = status; s.result = s.it;
}

// System q's finalize: its partial rows [rng[q], rng[q+1]) summed in a fixed order by
// the workgroup, then its scalar step in thread 0 (the same recurrences as the
// single-system path).
template <int OP, int NV, int NGRP = 1>
__device__ __forceinline__ void bt_fin_sys(int q, const double *part, const int *rng, BtSys *sy, int l, double *sh,
                                           bool valid = true) {
    const bool live = valid && (OP == BT_RESID || !sy[q].done);
    const int lt = phase_tid() & 255;
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) acc[i] = 0.0;
    if (live)
        for (int r = rng[q] + lt; r < rng[q + 1]; r += 256)
#pragma unroll
            for (int i = 0; i < NV; i++) acc[i] += part[(size_t)r * NV + i];
    group_sum<NV, NGRP>(acc, sh);
    if (lt != 0 || !live) return;
    BtSys &S = sy[q];
    if constexpr (OP == BT_RESID) {                         // FP64 true residual: ||r||, ||b||, ratio
        S.err = sqrt(acc[0]);
        S.fnb = sqrt(acc[1]);
        S.rel = acc[1] > 0.0 ? sqrt(acc[0] / acc[1]) : 0.0;
        return;
    }
    if constexpr (OP == BT_INIT) {
        S.fnb = sqrt(acc[1]);
        if (!(S.fnb > 0.0)) { bt_stop(S, ST_BREAKDOWN); return; }
        if (acc[0] == 0.0) { S.done = 1; S.status = ST_CONVERGED; S.result = 0; return; }
        S.alpha = 0.0; S.beta = 0.0; S.g0 = acc[0];
        return;
    }
    if constexpr (OP == BT_ALPHA) {
        if (bad(acc[0])) { bt_stop(S, ST_BREAKDOWN); return; }
        S.alpha = S.g0 / acc[0];
        return;
    }
    if constexpr (OP == BT_GAMMA) {
        if (bad(S.g0)) { bt_stop(S, ST_BREAKDOWN); return; }
        S.beta = -S.alpha * acc[0] / S.g0;
        S.g0 = acc[0];
        return;
    }
    if constexpr (OP == BT_MR) {
        // l from the Gram's size, NV = (l + 1)(l + 2) / 2, as a constant: the MR step's
        // arrays stay in registers
        constexpr int l = NV == 3 ? 1 : NV == 6 ? 2 : NV == 10 ? 3 : NV == 15 ? 4 : NV == 21 ? 5 : NV == 28 ? 6
                        : NV == 36 ? 7 : 8;
        static_assert((l + 1) * (l + 2) / 2 == NV, "MR Gram size");
        double G[MR_D][MR_D];
        gram_unpack(l, acc, G);
        if (G[0][0] == 0.0) { S.done = 1; S.status = ST_CONVERGED; S.result = S.it; return; }
        MrOut o;
        if (!mr_from_gram(l, G, o)) { bt_stop(S, ST_BREAKDOWN); return; }
        S.cx[0] = o.om[1];
        for (int k = 1; k < l; k++) {
            double t = 0.0;
            for (int i = k; i < l; i++) t += o.odd[i] * o.C[i][k];
            S.cx[k] = t;
        }
        for (int k = 1; k <= l; k++) {
            double t = 0.0;
            for (int i = k; i <= l; i++) t += o.od[i] * o.C[i][k];
            S.cr[k] = t;
            S.cp[k] = o.om[k];
        }
        S.oml = o.om[l];
        return;
    }
    // BT_UPD: error = ||r0|| / ||b|| (:122-125), then the next first beta (:57-62)
    if constexpr (OP == BT_UPD) {
        const double err = sqrt(acc[0]) / S.fnb;
        S.err = err;
        const int it = S.it;
        S.it = it + 1;
        if (err < S.tol) { S.done = 1; S.status = ST_CONVERGED; S.result = it; return; }
        if (!(err == err) || err > 1.0e300) { S.done = 1; S.status = ST_BREAKDOWN; S.result = it; return; }
        if (it + 1 >= S.max_it) { S.done = 1; S.status = ST_MAXIT; S.result = it; return; }
        const double g0 = -S.oml * S.g0;
        if (bad(g0)) { bt_stop(S, ST_BREAKDOWN); return; }
        S.beta = -S.alpha * acc[1] / g0;
        S.g0 = acc[1];
    }
}

// One workgroup per system.
template <int OP, int NV>
__global__ __launch_bounds__(256) void k_bt_fin(const double *part, const int *rng, BtSys *sy, int l) {
    __shared__ double sh[NV * 4];
    bt_fin_sys<OP, NV>(blockIdx.x, part, rng, sy, l, sh);
}

// ------------------------------------------------------- persistent solve
//
// The whole batched BiCGSTAB(L) -- initial residual and every outer iteration of
// every system -- as ONE launch: a resident grid (cooperative launch, so every
// workgroup is on the device at once) steps through the phases of the graph path,
// each phase spread over the grid, with a grid barrier between phases. Units, tiles
// and systems are dealt to workgroups XCD by XCD (workgroup g runs on XCD g % 8 and
// takes the g / 8-th share of that XCD's contiguous range), so a system's units share
// one L2 as in the per-phase launches. The loop ends when no system is active or
// after max_it iterations, the same test k_bt_alldone makes.
//
// Grid barrier, by XCD. The workgroups of one XCD (workgroup g runs on XCD g % 8,
// checked against the XCC_ID hardware register at entry) arrive, once their stores
// have completed (they sit in the XCD's shared L2), on a sub-group counter (8
// workgroups each), whose last arriver arrives on the XCD's monotone counter. The
// XCD's last arriver is its leader for this barrier: an agent-scope
// fence writes the XCD's L2 back, it arrives on the global counter (barrier k
// completes at 8 k leader arrivals; the last publishes generation k), waits for it,
// invalidates the XCD's L2 with a second agent-scope fence and publishes the XCD's
// generation; the sub-group leaders wait for that and publish their sub-group's, the
// others wait for their sub-group's; all make an acquire-only fence. So each
// barrier writes each XCD's L2 back once, as a kernel boundary does, instead of once
// per workgroup. If any workgroup finds itself on another XCD than g % 8, every barrier
// falls back to agent-scope fences in every workgroup. All waits are bounded: a wait
// that exceeds its bound sets `abort`, every waiting workgroup leaves, and the host
// reports a device error, so every wave exits.
struct BtGrid {
    unsigned arrive, gen;
    int abort, iters;
    unsigned xarrive[BKH_NXCD], xgen[BKH_NXCD];
    unsigned sarrive[BKH_NXCD][8], sgen[BKH_NXCD][8];   // sub-groups of an XCD's workgroups
    int bad_map, pad;
    unsigned qctr[3][BKH_NXCD];   // per-XCD item queues of the chunked persistent kernel
};

#define BT_SPIN_MAX (1u << 24)   // ~1 s of 128-cycle sleeps

// Waits until *w >= target; false (and `abort` set) if more than spin_max polls pass
// or another workgroup aborted.
__device__ __forceinline__ bool bt_wait(BtGrid *g, unsigned *w, unsigned target, unsigned spin_max) {
    unsigned spins = 0;
    while (__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (__hip_atomic_load(&g->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || ++spins > spin_max) {
            __hip_atomic_store(&g->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    return true;
}

// XCC this wave runs on (hardware register XCC_ID, bits 3:0).
__device__ __forceinline__ int bt_xcc_id() {
    return __builtin_amdgcn_s_getreg((3 << 11) | 20) & 0xF;
}

// Barrier sequence state of a workgroup: barriers passed, and the global and per-XCD
// arrival counts at the end of the last one (every workgroup passes the same
// sequence of barriers in the same modes, so all agree on them).
struct BtSeq {
    unsigned gen = 0, arrived = 0, xarrived = 0, sarrived = 0;
};

__device__ __forceinline__ bool bt_grid_sync(BtGrid *g, unsigned nwg, BtSeq &q, int *flag, bool hier,
                                             unsigned spin_max) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned target = q.gen + 1;
        if (!hier) {
            __threadfence();
            if (atomicAdd(&g->arrive, 1u) == q.arrived + nwg - 1)
                __hip_atomic_store(&g->gen, target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            else
                bt_wait(g, &g->gen, target, spin_max);
            __threadfence();
        } else {
            // arrival in three levels: a sub-group of up to 8 of the XCD's workgroups,
            // the XCD's sub-group leaders, the 8 XCD leaders (64 arrivals on one counter
            // serialise: 24-26 us per barrier against 11-13 us this way, measured alone)
            const int x = (int)blockIdx.x % BKH_NXCD, w = (int)blockIdx.x / BKH_NXCD, sg = w % 8;
            const unsigned nx = nwg / BKH_NXCD, nsg = nx < 8 ? nx : 8, nin = (nx - (unsigned)sg + 7) / 8;
            __builtin_amdgcn_s_waitcnt(0);                 // this workgroup's stores are in L2
            if (atomicAdd(&g->sarrive[x][sg], 1u) == q.sarrived + nin - 1) {
                if (atomicAdd(&g->xarrive[x], 1u) == q.xarrived + nsg - 1) {
                    __threadfence();                        // write the XCD's L2 back
                    if (atomicAdd(&g->arrive, 1u) == q.arrived + BKH_NXCD - 1)
                        __hip_atomic_store(&g->gen, target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                    else
                        bt_wait(g, &g->gen, target, spin_max);
                    __threadfence();                        // invalidate the XCD's L2 (and L1)
                    __hip_atomic_store(&g->xgen[x], target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                } else {
                    bt_wait(g, &g->xgen[x], target, spin_max);
                }
                __hip_atomic_store(&g->sgen[x][sg], target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                bt_wait(g, &g->sgen[x][sg], target, spin_max);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");       // acquire only: no L2 write-back
        }
        *flag = __hip_atomic_load(&g->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    q.gen++;
    if (hier) {
        const unsigned nx = nwg / BKH_NXCD, sg = (blockIdx.x / BKH_NXCD) % 8;
        q.arrived += BKH_NXCD;
        q.xarrived += nx < 8 ? nx : 8;
        q.sarrived += (nx - sg + 7) / 8;
    } else {
        q.arrived += nwg;
    }
    return *flag == 0;
}

// Work item k of `count` for workgroup g of G (G a multiple of BKH_NXCD): the XCD's
// contiguous share of [0, count), strided by its G / BKH_NXCD workgroups.
#define BT_DEAL(k, count)                                                                    \
    for (int x_ = (int)blockIdx.x % BKH_NXCD, per_ = ((count) + BKH_NXCD - 1) / BKH_NXCD,      \
             k = x_ * per_ + (int)blockIdx.x / BKH_NXCD,                                     \
             end_ = min((count), (x_ + 1) * per_);                                            \
         k < end_; k += (int)gridDim.x / BKH_NXCD)

// Next SpMV item for this workgroup, or -1: the XCD's contiguous share of [0, count),
// handed out one item at a time from the XCD's queue counter of the current phase
// (items differ in cost by an order of magnitude; a static deal leaves workgroups idle
// at the barrier). `slot` broadcasts thread 0's draw.
__device__ __forceinline__ int bt_grab(BtGrid *g, unsigned gen, int count, int *slot) {
    const int x = (int)blockIdx.x % BKH_NXCD, per = (count + BKH_NXCD - 1) / BKH_NXCD;
    const int end = min(count, (x + 1) * per);
    __syncthreads();                                   // the previous draw has been read
    if (threadIdx.x == 0) {
        const unsigned k = __hip_atomic_fetch_add(&g->qctr[gen % 3][x], 1u, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
        *slot = x * per + (int)k < end ? x * per + (int)k : -1;
    }
    __syncthreads();
    return *slot;
}

template <typename VT, typename T>
struct BtPArgs {
    const int *rp, *ci, *wblk, *usys, *ub, *trow, *tsys, *tb;
    const VT *val;
    const T *b, *rs_in;
    T *x, *R, *P, *rs;
    BtSys *sy;
    double *part;
    BtGrid *grid;
    long long N;
    int nunit, ntile, ns, max_it;
    unsigned spin_max;       // poll bound of every barrier wait
};

template <int L, typename VT, typename T>
__global__ __launch_bounds__(256) void k_bt_persist(BtPArgs<VT, T> a) {
    constexpr int NG = (L + 1) * (L + 2) / 2;
    __shared__ T prod_all[4][SPMV_WCAP];
    __shared__ double sh[NG * 4];
    __shared__ int flag, active;
    const unsigned nwg = gridDim.x;
    const long long N = a.N;
    BtSeq seq;
    Epi<0, T> e0;
    Epi<1, T> e1;
    Epi<NG, T> eg;
    e1.w[0] = a.rs;
    e1.a[0] = 0;
    e1.b[0] = EPI_Y;
    for (int k = 0; k < L; k++) eg.w[k] = a.R + k * N;
    {
        int idx = 0;
        for (int p = 0; p <= L; p++)
            for (int q = p; q <= L; q++) {
                eg.a[idx] = (unsigned char)(p == L ? EPI_Y : p);
                eg.b[idx] = (unsigned char)(q == L ? EPI_Y : q);
                idx++;
            }
    }
    if (threadIdx.x == 0 && bt_xcc_id() != (int)blockIdx.x % BKH_NXCD) atomicOr(&a.grid->bad_map, 1);
    bool hier = false;
#define BT_SYNC() if (!bt_grid_sync(a.grid, nwg, seq, &flag, hier, a.spin_max)) return
    // r0 = b - A x, p0 = rs = r0, then each system's initial scalars
    BT_DEAL(u, a.nunit) bt_spmv_unit<VT, T, 0>(u, a.rp, a.ci, a.val, a.wblk, a.usys, a.sy, a.x, a.R, e0, nullptr,
                                               prod_all, sh);
    BT_SYNC();                                   // full fences: publishes bad_map too
    hier = __hip_atomic_load(&a.grid->bad_map, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    BT_DEAL(t, a.ntile) bt_init_tile<T>(t, a.trow, a.tsys, a.b, a.R, a.P, a.rs, a.sy, a.part, sh);
    BT_SYNC();
    BT_DEAL(q, a.ns) bt_fin_sys<BT_INIT, 2>(q, a.part, a.tb, a.sy, L, sh);
    BT_SYNC();
    int it = 0;
    for (; it < a.max_it; it++) {
        for (int j = 0; j < L; j++) {
            BT_DEAL(t, a.ntile) bt_p_tile<T, 1>(t, a.trow, a.tsys, N, j, a.P, a.R, a.sy);
            BT_SYNC();
            BT_DEAL(u, a.nunit) bt_spmv_unit<VT, T, 1>(u, a.rp, a.ci, a.val, a.wblk, a.usys, a.sy, a.P + j * N,
                                                       a.P + (j + 1) * N, e1, a.part, prod_all, sh);
            BT_SYNC();
            BT_DEAL(q, a.ns) bt_fin_sys<BT_ALPHA, 1>(q, a.part, a.ub, a.sy, L, sh);
            BT_SYNC();
            BT_DEAL(t, a.ntile) bt_r_tile<T, 1>(t, a.trow, a.tsys, N, j, a.P, a.R, a.x, a.sy);
            BT_SYNC();
            if (j < L - 1) {
                BT_DEAL(u, a.nunit) bt_spmv_unit<VT, T, 1>(u, a.rp, a.ci, a.val, a.wblk, a.usys, a.sy, a.R + j * N,
                                                           a.R + (j + 1) * N, e1, a.part, prod_all, sh);
                BT_SYNC();
                BT_DEAL(q, a.ns) bt_fin_sys<BT_GAMMA, 1>(q, a.part, a.ub, a.sy, L, sh);
            } else {
                BT_DEAL(u, a.nunit) bt_spmv_unit<VT, T, NG>(u, a.rp, a.ci, a.val, a.wblk, a.usys, a.sy,
                                                            a.R + (L - 1) * N, a.R + L * N, eg, a.part, prod_all, sh);
                BT_SYNC();
                BT_DEAL(q, a.ns) bt_fin_sys<BT_MR, NG>(q, a.part, a.ub, a.sy, L, sh);
            }
            BT_SYNC();
        }
        BT_DEAL(t, a.ntile) bt_upd_tile<T, 1, 1>(t, a.trow, a.tsys, N, L, a.P, a.R, a.x, a.rs, a.sy, a.part, sh);
        BT_SYNC();
        BT_DEAL(q, a.ns) bt_fin_sys<BT_UPD, 2>(q, a.part, a.tb, a.sy, L, sh);
        BT_SYNC();
        // every workgroup reads the same settled flags, so all leave together
        if (threadIdx.x == 0) active = 0;
        __syncthreads();
        int mine = 0;
        for (int q = threadIdx.x; q < a.ns; q += 256) mine += a.sy[q].done ? 0 : 1;
        if (mine) atomicAdd(&active, mine);
        __syncthreads();
        if (active == 0) { it++; break; }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) a.grid->iters = it;
#undef BT_SYNC
}

// The persistent solve with the chunked SpMV: LC_NT-thread workgroups (two per CU), each
// SpMV item taken by a whole workgroup (bt_spmv_item, as k_bt_spmv_lc), tiles and
// systems by its four 256-thread groups (four per workgroup step, the same per-tile and
// per-system arithmetic as the 256-thread per-phase kernels); items are drawn from
// per-XCD queues (bt_grab), tiles and systems dealt as in k_bt_persist. x is bitwise the graph
// path's with the chunked SpMV. LDS: the x chunk and the products only; the groups'
// partial sums use the product buffer, the barrier's flags the x chunk's first words
// (each is free whenever the other use runs).

template <typename T>
__device__ __forceinline__ Epi<1, T> epi_dot(const T *w) {
    Epi<1, T> e;
    e.w[0] = w;
    e.a[0] = 0;
    e.b[0] = EPI_Y;
    return e;
}
// The MR step's Gram entries (r_p, r_q), p <= q <= L, r_L = y.
template <int L, typename T>
__device__ __forceinline__ Epi<(L + 1) * (L + 2) / 2, T> epi_gram(T *R, long long N) {
    Epi<(L + 1) * (L + 2) / 2, T> e;
    for (int k = 0; k < L; k++) e.w[k] = R + k * N;
    int idx = 0;
    for (int p = 0; p <= L; p++)
        for (int q = p; q <= L; q++) {
            e.a[idx] = (unsigned char)(p == L ? EPI_Y : p);
            e.b[idx] = (unsigned char)(q == L ? EPI_Y : q);
            idx++;
        }
    return e;
}

template <int L, typename VT, typename T>
__global__ __launch_bounds__(LC_NT) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_bt_persist_lc(
    BtPArgs<VT, T> A, LcArgs<VT> la, const int *ib, int nitem) {
    constexpr int NG = (L + 1) * (L + 2) / 2, BLK = lc_blk<T>(), NGRP = LC_NT / 256;
    static_assert(NGRP * NG * 4 * sizeof(double) <= BLK * sizeof(T), "group sums fit in prod");
    __shared__ T xs[LC_CH];
    __shared__ __attribute__((aligned(32))) T prod[BLK];
    double *sh = (double *)prod;
    int *flag = (int *)xs, *active = (int *)xs + 1, *slot = (int *)xs + 2;
    const unsigned nwg = gridDim.x;
    BtSeq seq;
    if (threadIdx.x == 0 && bt_xcc_id() != (int)blockIdx.x % BKH_NXCD) atomicOr(&A.grid->bad_map, 1);
    bool hier = false;
    // after barrier g, the queue of phase g + 1 (last drawn from in phase g - 2) is reset
#define BT_SYNC()                                                                                     \
    if (!bt_grid_sync(A.grid, nwg, seq, flag, hier, A.spin_max)) return;                              \
    if (threadIdx.x == 0 && blockIdx.x < BKH_NXCD)                                                    \
    __hip_atomic_store(&A.grid->qctr[(seq.gen + 1) % 3][blockIdx.x], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
#define BT_ITEMS(it) for (int it = bt_grab(A.grid, seq.gen, nitem, slot); it >= 0; it = bt_grab(A.grid, seq.gen, nitem, slot))
#define BT_TILES(t) BT_DEAL(k_, (A.ntile + NGRP - 1) / NGRP) for (int t = k_ * NGRP + (phase_tid() >> 8), once_ = 1; once_; once_ = 0)
#define BT_SYSTEMS(q) BT_DEAL(k_, (A.ns + NGRP - 1) / NGRP) for (int q = k_ * NGRP + (phase_tid() >> 8), once_ = 1; once_; once_ = 0)
    {
        Epi<0, T> e0;
        BT_ITEMS(it) bt_spmv_item<VT, T, 0, BLK>(it, la, A.sy, A.x, A.R, e0, nullptr, xs, prod);
    }
    BT_SYNC();
    hier = __hip_atomic_load(&A.grid->bad_map, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    BT_TILES(t) bt_init_tile<T, NGRP>(t, A.trow, A.tsys, A.b, A.R, A.P, A.rs, A.sy, A.part, sh, t < A.ntile);
    BT_SYNC();
    BT_SYSTEMS(q) bt_fin_sys<BT_INIT, 2, NGRP>(q, A.part, A.tb, A.sy, L, sh, q < A.ns);
    BT_SYNC();
    int it = 0;
    for (; it < A.max_it; it++) {
        for (int j = 0; j < L; j++) {
            BT_TILES(t) if (t < A.ntile) bt_p_tile<T, 1>(t, A.trow, A.tsys, A.N, j, A.P, A.R, A.sy);
            BT_SYNC();
            BT_ITEMS(u) bt_spmv_item<VT, T, 1, BLK>(u, la, A.sy, A.P + j * A.N, A.P + (j + 1) * A.N,
                                                              epi_dot<T>(A.rs), A.part, xs, prod);
            BT_SYNC();
            BT_SYSTEMS(q) bt_fin_sys<BT_ALPHA, 1, NGRP>(q, A.part, ib, A.sy, L, sh, q < A.ns);
            BT_SYNC();
            BT_TILES(t) if (t < A.ntile) bt_r_tile<T, 1>(t, A.trow, A.tsys, A.N, j, A.P, A.R, A.x, A.sy);
            BT_SYNC();
            if (j < L - 1) {
                BT_ITEMS(u) bt_spmv_item<VT, T, 1, BLK>(u, la, A.sy, A.R + j * A.N,
                                                                  A.R + (j + 1) * A.N, epi_dot<T>(A.rs), A.part,
                                                                  xs, prod);
                BT_SYNC();
                BT_SYSTEMS(q) bt_fin_sys<BT_GAMMA, 1, NGRP>(q, A.part, ib, A.sy, L, sh, q < A.ns);
            } else {
                BT_ITEMS(u) bt_spmv_item<VT, T, NG, BLK>(u, la, A.sy, A.R + (L - 1) * A.N,
                                                                   A.R + L * A.N, epi_gram<L, T>(A.R, A.N),
                                                                   A.part, xs, prod);
                BT_SYNC();
                BT_SYSTEMS(q) bt_fin_sys<BT_MR, NG, NGRP>(q, A.part, ib, A.sy, L, sh, q < A.ns);
            }
            BT_SYNC();
        }
        BT_TILES(t) bt_upd_tile<T, NGRP, 1>(t, A.trow, A.tsys, A.N, L, A.P, A.R, A.x, A.rs, A.sy, A.part, sh,
                                         t < A.ntile);
        BT_SYNC();
        BT_SYSTEMS(q) bt_fin_sys<BT_UPD, 2, NGRP>(q, A.part, A.tb, A.sy, L, sh, q < A.ns);
        BT_SYNC();
        // every workgroup reads the same settled flags, so all leave together
        if (threadIdx.x == 0) *active = 0;
        __syncthreads();
        int mine = 0;
        for (int q = threadIdx.x; q < A.ns; q += LC_NT) mine += A.sy[q].done ? 0 : 1;
        if (mine) atomicAdd(active, mine);
        __syncthreads();
        const bool stop = *active == 0;
        __syncthreads();                           // *active read before the next phase's LDS stores
        if (stop) { it++; break; }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) A.grid->iters = it;
#undef BT_ITEMS
#undef BT_TILES
#undef BT_SYSTEMS
#undef BT_SYNC
}

// Stops the graph once no system is active.
__global__ __launch_bounds__(256) void k_bt_alldone(const BtSys *sy, int ns, Ctl *c) {
    __shared__ int cnt[256];
    int a = 0;
    for (int s = threadIdx.x; s < ns; s += 256) a += sy[s].done ? 0 : 1;
    cnt[threadIdx.x] = a;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) cnt[threadIdx.x] += cnt[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        c->it += 1;
        if (cnt[0] == 0) { c->done = 1; c->status = ST_CONVERGED; }
        else if (c->it >= c->max_it) { c->done = 1; c->status = ST_MAXIT; }
    }
}

// Per-system control at the start of a (pass of a) solve.
__global__ void k_bt_sys_init(BtSys *sy, int ns, const double *tol_each, double tol, const int *frozen,
                              int max_it) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    BtSys &S = sy[s];
    S.tol = tol_each ? tol_each[s] : tol;
    S.max_it = max_it;
    S.it = 0; S.result = 0; S.err = 0.0;
    S.alpha = 0.0; S.beta = 0.0; S.g0 = 0.0; S.oml = 1.0;
    const bool fz = frozen && frozen[s];
    S.done = (fz || max_it <= 0) ? 1 : 0;
    S.status = fz ? ST_CONVERGED : (max_it <= 0 ? ST_MAXIT : ST_RUNNING);
}

// r = b - r (A x in r), partials {r.r, b.b} (FP64 true residual of every system).
__global__ __launch_bounds__(256) void k_bt_resid(const int *trow, const int *tsys, const double *b, double *r,
                                                  double *part) {
    double v[2] = {0.0, 0.0};
    TILE_ROWS(i, blockIdx.x) {
        const double ri = b[i] - r[i];
        r[i] = ri;
        v[0] += ri * ri;
        v[1] += b[i] * b[i];
    }
    (void)sys;
    __shared__ double sh[8];
    unit_partial<2>(v, part, blockIdx.x, sh);
}

// r32 = r64 (scaled by 1/||r_s|| so every inner system starts at unit norm), d = 0.
__global__ __launch_bounds__(256) void k_bt_ir_prep(const int *trow, const int *tsys, const double *r64,
                                                    float *r32, float *d, const double *scale) {
    TILE_ROWS(i, blockIdx.x) {
        r32[i] = (float)(r64[i] * scale[sys]);
        d[i] = 0.0f;
    }
}

// x += d / scale_s (FP64)
__global__ __launch_bounds__(256) void k_bt_ir_update(const int *trow, const int *tsys, double *x, const float *d,
                                                      const double *scale) {
    TILE_ROWS(i, blockIdx.x) x[i] += (double)d[i] / scale[sys];
}

// --------------------------------------------------------------- building

static void bt_free(bkh_batch *B) {
    hipFree(B->rp); hipFree(B->ci); hipFree(B->val); hipFree(B->val32);
    hipFree(B->items); hipFree(B->lval); hipFree(B->lval32); hipFree(B->lcol); hipFree(B->lrp); hipFree(B->lcb);
    hipFree(B->ib);
    hipFree(B->wblk); hipFree(B->usys); hipFree(B->ub);
    hipFree(B->trow); hipFree(B->tsys); hipFree(B->tb);
}

static void bt_free_layout(bkh_batch *B) {
    bkh_graph_epoch_bump();
    hipFree(B->wblk); hipFree(B->usys); hipFree(B->ub);
    hipFree(B->trow); hipFree(B->tsys); hipFree(B->tb);
    B->wblk = B->usys = B->ub = B->trow = B->tsys = B->tb = nullptr;
}

// SpMV units and vector tiles for the given caps.
static int bt_layout(bkh_batch *B, int rows_cap, int nnz_cap, int tile_rows) {
    std::vector<int> wb, us, ub, tr, ts, tb;
    wb.push_back(0);
    ub.push_back(0);
    tb.push_back(0);
    tr.push_back(0);
    for (int s = 0; s < B->ns; s++) {
        const int a = (int)B->off[s], e = (int)B->off[s + 1];
        int nb = 0, r = a;
        while (r < e) {
            int start = r, nz = 0;
            while (r < e && r - start < rows_cap) {
                const int l = B->rp_h[r + 1] - B->rp_h[r];
                if (nz + l > nnz_cap && r > start) break;
                nz += l;
                r++;
            }
            wb.push_back(r);
            nb++;
        }
        while (nb % 4) { wb.push_back(e); nb++; }
        for (int q = 0; q < nb / 4; q++) us.push_back(s);
        ub.push_back((int)us.size());
        for (int r0 = a; r0 < e; r0 += tile_rows) {
            tr.push_back(std::min(e, r0 + tile_rows));
            ts.push_back(s);
        }
        tb.push_back((int)ts.size());
    }
    bt_free_layout(B);
    B->nunit = (int)us.size();
    B->ntile = (int)ts.size();
    bool ok = bkh_dmalloc(&B->wblk, sizeof(int) * wb.size()) == hipSuccess &&
              bkh_dmalloc(&B->usys, sizeof(int) * std::max<size_t>(us.size(), 1)) == hipSuccess &&
              bkh_dmalloc(&B->ub, sizeof(int) * ub.size()) == hipSuccess &&
              bkh_dmalloc(&B->trow, sizeof(int) * tr.size()) == hipSuccess &&
              bkh_dmalloc(&B->tsys, sizeof(int) * std::max<size_t>(ts.size(), 1)) == hipSuccess &&
              bkh_dmalloc(&B->tb, sizeof(int) * tb.size()) == hipSuccess;
    ok = ok && hipMemcpy(B->wblk, wb.data(), sizeof(int) * wb.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(B->usys, us.data(), sizeof(int) * us.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(B->ub, ub.data(), sizeof(int) * ub.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(B->trow, tr.data(), sizeof(int) * tr.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(B->tsys, ts.data(), sizeof(int) * ts.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(B->tb, tb.data(), sizeof(int) * tb.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) { bkh_set_error("batch layout upload failed"); return BKH_DEVICE_ERROR; }
    B->rows_cap = rows_cap;
    B->nnz_cap = nnz_cap;
    B->tile_rows = tile_rows;
    return 0;
}

// The column-chunked copy (see k_bt_spmv_lc) when every row's columns ascend; the
// SpMV then runs the chunked kernel by default. Returns 0 or a device error.
static int bt_build_lc(bkh_batch *B, const std::vector<int> &ci, const std::vector<double> &v, int want_fp32) {
    for (long long r = 0; r < B->N; r++)
        for (int k = B->rp_h[r] + 1; k < B->rp_h[r + 1]; k++)
            if (ci[k] <= ci[k - 1]) return 0;                   // a row not in ascending order
    std::vector<LcItem> items;
    std::vector<int> cbase, ib{0};
    std::vector<unsigned short> lrp;
    std::vector<double> lv;
    std::vector<unsigned short> lc;
    lv.reserve((size_t)B->nnz + 4);
    lc.reserve((size_t)B->nnz + 4);
    for (int s = 0; s < B->ns; s++) {
        const int xo = (int)B->off[s], n = (int)(B->off[s + 1] - B->off[s]);
        const int nc = (n + LC_CH - 1) / LC_CH;
        std::vector<int> cnt(nc);
        for (int r0 = 0; r0 < n;) {
            // the most rows (<= LC_ROWS) whose entries in every chunk fit 16-bit offsets
            int R = std::min(LC_ROWS, n - r0);
            for (;;) {
                std::fill(cnt.begin(), cnt.end(), 0);
                for (int k = B->rp_h[xo + r0]; k < B->rp_h[xo + r0 + R]; k++) cnt[(ci[k] - xo) / LC_CH]++;
                if (*std::max_element(cnt.begin(), cnt.end()) <= 65535 || R == 1) break;
                R = (R + 1) / 2;
            }
            LcItem it;
            it.sys = s; it.xoff = xo; it.r0 = xo + r0; it.R = R; it.nc = nc;
            it.rpb = (int)lrp.size(); it.ns = n; it.cbi = (int)cbase.size();
            std::vector<int> cur(R);
            for (int r = 0; r < R; r++) cur[r] = B->rp_h[it.r0 + r];
            for (int c = 0; c < nc; c++) {
                const int hi = (c + 1) * LC_CH;
                const int base = (int)lv.size();                   // a multiple of 8
                cbase.push_back(base);
                for (int r = 0; r < R; r++) {
                    lrp.push_back((unsigned short)(lv.size() - base));
                    const int end = B->rp_h[it.r0 + r + 1];
                    while (cur[r] < end && ci[cur[r]] - xo < hi) {
                        lv.push_back(v[cur[r]]);
                        lc.push_back((unsigned short)(ci[cur[r]] - xo - c * LC_CH));
                        cur[r]++;
                    }
                }
                lrp.push_back((unsigned short)(lv.size() - base));
                while (lv.size() % 8) { lv.push_back(0.0); lc.push_back(0); }   // next chunk 8-aligned
            }
            items.push_back(it);
            r0 += R;
        }
        ib.push_back((int)items.size());
    }
    if (lv.size() >= (size_t)INT_MAX - 16) return 0;          // entry indices stay int32
    const size_t ne = lv.size() + 8;
    lv.resize(ne, 0.0);
    lc.resize(ne, 0);
    bool ok = bkh_dmalloc(&B->items, sizeof(LcItem) * items.size()) == hipSuccess &&
              bkh_dmalloc(&B->lval, sizeof(double) * ne) == hipSuccess &&
              bkh_dmalloc(&B->lcol, sizeof(unsigned short) * ne) == hipSuccess &&
              bkh_dmalloc(&B->lrp, sizeof(unsigned short) * lrp.size()) == hipSuccess &&
              bkh_dmalloc(&B->lcb, sizeof(int) * cbase.size()) == hipSuccess &&
              bkh_dmalloc(&B->ib, sizeof(int) * ib.size()) == hipSuccess;
    if (ok && want_fp32) ok = bkh_dmalloc(&B->lval32, sizeof(float) * ne) == hipSuccess;
    ok = ok && hipMemcpy(B->items, items.data(), sizeof(LcItem) * items.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(B->lval, lv.data(), sizeof(double) * ne, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(B->lcol, lc.data(), sizeof(unsigned short) * ne, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(B->lrp, lrp.data(), sizeof(unsigned short) * lrp.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(B->lcb, cbase.data(), sizeof(int) * cbase.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(B->ib, ib.data(), sizeof(int) * ib.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (ok && want_fp32) {
        std::vector<float> lv32(ne);
        for (size_t q = 0; q < ne; q++) lv32[q] = (float)lv[q];
        ok = hipMemcpy(B->lval32, lv32.data(), sizeof(float) * ne, hipMemcpyHostToDevice) == hipSuccess;
    }
    if (!ok) { bkh_set_error("batch chunked copy upload failed"); return BKH_DEVICE_ERROR; }
    B->nitem = (int)items.size();
    B->lc_ok = 1;
    B->use_lc = 1;
    return 0;
}

extern "C" bkh_batch *bkh_batch_create(int ns, const int *n_each, const int *const *row_ptr,
                                       const int *const *col, const double *const *val, int want_fp32) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return nullptr;
    if (ns < 1 || !n_each || !row_ptr || !col || !val) { bkh_set_error("bad batch arguments"); return nullptr; }
    bkh_batch *B = new bkh_batch;
    B->ns = ns;
    B->off.assign(ns + 1, 0);
    for (int s = 0; s < ns; s++) {
        if (n_each[s] < 1 || !row_ptr[s] || row_ptr[s][0] != 0) { bkh_set_error("system %d: bad CSR", s); delete B; return nullptr; }
        B->off[s + 1] = B->off[s] + n_each[s];
        B->nnz += row_ptr[s][n_each[s]];
    }
    B->N = B->off[ns];
    if (B->N >= (1LL << 31) || B->nnz >= (1LL << 31) - 8) { bkh_set_error("batch exceeds int32 indexing"); delete B; return nullptr; }
    B->rp_h.assign(B->N + 1, 0);
    std::vector<int> ci((size_t)B->nnz + 4, 0);
    std::vector<double> v((size_t)B->nnz + 4, 0.0);
    long long k = 0;
    for (int s = 0; s < ns; s++) {
        const long long o = B->off[s];
        const int n = n_each[s];
        for (int i = 0; i < n; i++) {
            const int a = row_ptr[s][i], e = row_ptr[s][i + 1];
            if (e < a) { bkh_set_error("system %d: row pointers decrease", s); delete B; return nullptr; }
            B->max_row = std::max(B->max_row, e - a);
            for (int q = a; q < e; q++) {
                const int c = col[s][q];
                if (c < 0 || c >= n) { bkh_set_error("system %d: column out of range", s); delete B; return nullptr; }
                ci[k] = (int)(c + o);
                v[k] = val[s][q];
                k++;
            }
            B->rp_h[o + i + 1] = (int)k;
        }
    }
    if (B->max_row > SPMV_WCAP) { bkh_set_error("batched SpMV supports rows of <= %d nonzeros", SPMV_WCAP); delete B; return nullptr; }
    const size_t nz = (size_t)B->nnz + 4;
    bool ok = bkh_dmalloc(&B->rp, sizeof(int) * (B->N + 1)) == hipSuccess &&
              bkh_dmalloc(&B->ci, sizeof(int) * nz) == hipSuccess &&
              bkh_dmalloc(&B->val, sizeof(double) * nz) == hipSuccess;
    if (ok && want_fp32) ok = bkh_dmalloc(&B->val32, sizeof(float) * nz) == hipSuccess;
    ok = ok && hipMemcpy(B->rp, B->rp_h.data(), sizeof(int) * (B->N + 1), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(B->ci, ci.data(), sizeof(int) * nz, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(B->val, v.data(), sizeof(double) * nz, hipMemcpyHostToDevice) == hipSuccess;
    if (ok && want_fp32) {
        std::vector<float> v32(nz);
        for (size_t q = 0; q < nz; q++) v32[q] = (float)v[q];
        ok = hipMemcpy(B->val32, v32.data(), sizeof(float) * nz, hipMemcpyHostToDevice) == hipSuccess;
    }
    if (!ok || bt_layout(B, SPMV_WROWS, SPMV_WCAP, 1024)) {
        if (!ok) bkh_set_error("batch upload failed");
        bt_free(B);
        delete B;
        return nullptr;
    }
    if (bt_build_lc(B, ci, v, want_fp32)) {
        bt_free(B);
        delete B;
        return nullptr;
    }
    return B;
}

extern "C" void bkh_batch_destroy(bkh_batch *B) {
    if (!B) return;
    bkh_graph_epoch_bump();
    bt_free(B);
    delete B;
}

extern "C" long long bkh_batch_rows(const bkh_batch *B) { return B ? B->N : 0; }
extern "C" long long bkh_batch_nnz(const bkh_batch *B) { return B ? B->nnz : 0; }
extern "C" long long bkh_batch_offset(const bkh_batch *B, int s) {
    return (B && s >= 0 && s <= B->ns) ? B->off[s] : -1;
}
// 1 selects the column-chunked SpMV (BKH_BAD_ARGUMENT when the batch has no chunked
// copy: some row's columns do not ascend), 0 the CSR-stream units.
extern "C" int bkh_batch_set_spmv_kernel(bkh_batch *B, int chunked) {
    if (!B || (chunked && !B->lc_ok)) return BKH_BAD_ARGUMENT;
    B->use_lc = chunked ? 1 : 0;
    bkh_graph_epoch_bump();
    return 0;
}
extern "C" int bkh_batch_spmv_kernel(const bkh_batch *B) { return B ? B->use_lc : -1; }

extern "C" void bkh_batch_layout(const bkh_batch *B, int *rows_cap, int *nnz_cap, int *tile_rows) {
    if (!B) return;
    if (rows_cap) *rows_cap = B->rows_cap;
    if (nnz_cap) *nnz_cap = B->nnz_cap;
    if (tile_rows) *tile_rows = B->tile_rows;
}

// ----------------------------------------------------------------- launches

// The chunked copy's arrays, values in the precision asked for (FP64 or the FP32 copy).
template <typename VT>
static LcArgs<VT> lc_args(const bkh_batch *B) {
    LcArgs<VT> la;
    la.items = B->items;
    la.ev = (sizeof(VT) == sizeof(double)) ? (const VT *)B->lval : (const VT *)B->lval32;
    la.ec = B->lcol;
    la.erp = B->lrp;
    la.ecb = B->lcb;
    return la;
}
template <typename VT, typename T, int NV>
static void bt_spmv(hipStream_t s, const bkh_batch *B, const VT *val, const BtSys *sy, const T *x, T *y,
                    const Epi<NV, T> &e, double *part) {
    if (B->use_lc) {
        hipLaunchKernelGGL((k_bt_spmv_lc<VT, T, NV, lc_blk<T>()>), dim3(B->nitem), dim3(LC_NT), 0, s,
                           lc_args<VT>(B), sy, x, y, e, part);
        return;
    }
    hipLaunchKernelGGL((k_bt_spmv<VT, T, NV>), dim3(B->nunit), dim3(256), 0, s, B->rp, B->ci, val, B->wblk,
                       B->usys, sy, x, y, e, part);
}

template <int OP, int NV>
static void bt_fin(hipStream_t s, const bkh_batch *B, const double *part, bool units, BtSys *sy, int l) {
    hipLaunchKernelGGL((k_bt_fin<OP, NV>), dim3(B->ns), dim3(256), 0, s, part,
                       units ? (B->use_lc ? (const int *)B->ib : B->ub) : B->tb, sy, l);
}

template <typename VT, typename T>
static void bt_dot_spmv(hipStream_t s, const bkh_batch *B, const VT *val, BtSys *sy, const T *x, T *y,
                        const T *w, double *part) {
    Epi<1, T> e;
    memset(&e, 0, sizeof e);
    e.w[0] = w;
    e.a[0] = 0;
    e.b[0] = EPI_Y;
    bt_spmv<VT, T, 1>(s, B, val, sy, x, y, e, part);
}

template <int L, typename VT, typename T>
static void bt_gram_spmv(hipStream_t s, const bkh_batch *B, const VT *val, BtSys *sy, T *R, double *part) {
    constexpr int NV = (L + 1) * (L + 2) / 2;
    const long long N = B->N;
    Epi<NV, T> e;
    memset(&e, 0, sizeof e);
    for (int k = 0; k < L; k++) e.w[k] = R + k * N;
    int idx = 0;
    for (int a = 0; a <= L; a++)
        for (int b = a; b <= L; b++) {
            e.a[idx] = (unsigned char)(a == L ? EPI_Y : a);
            e.b[idx] = (unsigned char)(b == L ? EPI_Y : b);
            idx++;
        }
    bt_spmv<VT, T, NV>(s, B, val, sy, R + (L - 1) * N, R + L * N, e, part);
    bt_fin<BT_MR, NV>(s, B, part, true, sy, L);
}

static int g_bt_last_path = 0;      // 1 graph of per-phase launches, 2 persistent

// The whole solve as one cooperative launch of k_bt_persist (see there). The grid is
// what the device holds at once (occupancy x CUs), a multiple of the XCD count, and
// no larger than the largest work list.
template <int L, typename VT, typename T>
static int bt_core_persist(ThreadCtx *t, const bkh_batch *B, const VT *val, const T *b, T *x, T *R, T *P, T *rs,
                           BtSys *sy, double *part, BtGrid *grid, int max_it) {
    hipStream_t s = t->stream;
    const bool lc = B->use_lc != 0;
    static int G_res = [] {
        int dev = 0, ncu = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_bt_persist<L, VT, T>, 256, 0) != hipSuccess)
            return 0;
        return ncu * per;
    }();
    static int G_lc = [] {
        int dev = 0, ncu = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_bt_persist_lc<L, VT, T>, LC_NT, 0) != hipSuccess)
            return 0;
        return ncu * per;
    }();
    const int work = lc ? std::max(std::max(B->nitem, (B->ntile + 3) / 4), (B->ns + 3) / 4)
                        : std::max(std::max(B->nunit, B->ntile), B->ns);
    int G = std::min(lc ? G_lc : G_res, (work + BKH_NXCD - 1) / BKH_NXCD * BKH_NXCD);
    G -= G % BKH_NXCD;
    if (G < BKH_NXCD) { bkh_set_error("persistent batched solve: no resident grid (occupancy %d)", G_res); return BKH_DEVICE_ERROR; }
    HIP_OK(hipMemsetAsync(grid, 0, sizeof(BtGrid), s));
    BtPArgs<VT, T> a;
    a.rp = B->rp; a.ci = B->ci; a.wblk = B->wblk; a.usys = B->usys; a.ub = B->ub;
    a.trow = B->trow; a.tsys = B->tsys; a.tb = B->tb;
    a.val = val; a.b = b; a.rs_in = nullptr; a.x = x; a.R = R; a.P = P; a.rs = rs;
    a.sy = sy; a.part = part; a.grid = grid; a.N = B->N;
    a.nunit = B->nunit; a.ntile = B->ntile; a.ns = B->ns; a.max_it = max_it;
    // BKH_BT_SPIN lowers the barrier poll bound (0: a waiting workgroup gives up at once),
    // so the abort path can be exercised from the host; every wave still exits
    const char *sp = getenv("BKH_BT_SPIN");
    a.spin_max = (sp && *sp) ? (unsigned)strtoul(sp, nullptr, 10) : BT_SPIN_MAX;
    LcArgs<VT> la = lc_args<VT>(B);
    const int *ib = B->ib;
    int nitem = B->nitem;
    void *args[] = {&a};
    void *args_lc[] = {&a, &la, &ib, &nitem};
    const hipError_t e =
        lc ? hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&k_bt_persist_lc<L, VT, T>), dim3(G),
                                        dim3(LC_NT), args_lc, 0, s)
           : hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&k_bt_persist<L, VT, T>), dim3(G), dim3(256),
                                        args, 0, s);
    if (e != hipSuccess) { bkh_set_error("cooperative launch of the batched solve: %s", hipGetErrorString(e)); return BKH_DEVICE_ERROR; }
    BtGrid hg;
    HIP_OK(hipMemcpyAsync(&hg, grid, sizeof hg, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (hg.abort) { bkh_set_error("persistent batched solve: grid barrier wait exceeded its bound"); return BKH_DEVICE_ERROR; }
    g_bt_last_path = 2;
    return 0;
}

// One batched BiCGSTAB(L) solve (all systems) on vectors T; sy initialised by caller.
template <int L, typename VT, typename T>
static int bt_core(ThreadCtx *t, const bkh_batch *B, const VT *val, const T *b, T *x, T *R, T *P, T *rs,
                   BtSys *sy, Ctl *c, double *part, int max_it, int K, BtGrid *grid) {
    if (grid) return bt_core_persist<L, VT, T>(t, B, val, b, x, R, P, rs, sy, part, grid, max_it);
    g_bt_last_path = 1;
    hipStream_t s = t->stream;
    const long long N = B->N;
    if (ctl_init(t, c, 0.0, max_it)) return BKH_DEVICE_ERROR;
    Epi<0, T> e0;
    memset(&e0, 0, sizeof e0);
    bt_spmv<VT, T, 0>(s, B, val, sy, x, R, e0, nullptr);
    hipLaunchKernelGGL(k_bt_init<T>, dim3(B->ntile), dim3(256), 0, s, B->trow, B->tsys, N, b, R, P, rs,
                       (const BtSys *)sy, part);
    bt_fin<BT_INIT, 2>(s, B, part, false, sy, L);
    LAUNCH_OK();
    hipGraphExec_t ge = capture_cached(t, {0xBA00, (uintptr_t)B, (uintptr_t)val, (uintptr_t)b, (uintptr_t)x, (uintptr_t)R,
                                           (uintptr_t)P, (uintptr_t)rs, (uintptr_t)sy, (uintptr_t)c, (uintptr_t)part,
                                           (uintptr_t)L, (uintptr_t)sizeof(T)}, s, K, [&]() {
        for (int j = 0; j < L; j++) {
            hipLaunchKernelGGL(k_bt_p<T>, dim3(B->ntile), dim3(256), 0, s, B->trow, B->tsys, N, j, P, (const T *)R,
                               (const BtSys *)sy);
            bt_dot_spmv<VT, T>(s, B, val, sy, P + j * N, P + (j + 1) * N, rs, part);
            bt_fin<BT_ALPHA, 1>(s, B, part, true, sy, L);
            hipLaunchKernelGGL(k_bt_r<T>, dim3(B->ntile), dim3(256), 0, s, B->trow, B->tsys, N, j, (const T *)P, R,
                               x, (const BtSys *)sy);
            if (j < L - 1) {
                bt_dot_spmv<VT, T>(s, B, val, sy, R + j * N, R + (j + 1) * N, rs, part);
                bt_fin<BT_GAMMA, 1>(s, B, part, true, sy, L);
            } else {
                bt_gram_spmv<L, VT, T>(s, B, val, sy, R, part);
            }
        }
        hipLaunchKernelGGL(k_bt_upd<T>, dim3(B->ntile), dim3(256), 0, s, B->trow, B->tsys, N, L, P, R, x,
                           (const T *)rs, (const BtSys *)sy, part);
        bt_fin<BT_UPD, 2>(s, B, part, false, sy, L);
        hipLaunchKernelGGL(k_bt_alldone, dim3(1), dim3(256), 0, s, (const BtSys *)sy, B->ns, c);
    });
    if (!ge) { bkh_set_error("hipGraph capture of the batched iteration failed"); return BKH_DEVICE_ERROR; }
    int rc = run_graph_until_done(t, ge, c, max_it / K + 2);
    return rc;
}

// FP64 true residual of every system into sy[s].rel (r64 = b - A x).
static int bt_true_res(ThreadCtx *t, const bkh_batch *B, const double *b, const double *x, double *r64,
                       BtSys *sy, double *part) {
    hipStream_t s = t->stream;
    Epi<0, double> e0;
    memset(&e0, 0, sizeof e0);
    bt_spmv<double, double, 0>(s, B, B->val, nullptr, x, r64, e0, nullptr);
    hipLaunchKernelGGL(k_bt_resid, dim3(B->ntile), dim3(256), 0, s, B->trow, B->tsys, b, r64, part);
    bt_fin<BT_RESID, 2>(s, B, part, false, sy, 0);
    LAUNCH_OK();
    return 0;
}

template <int L>
static int bt_solve(ThreadCtx *t, const bkh_batch *B, const double *b, double *x, double tol, int max_iter,
                    const bkh_opts &o, int *ret, double *true_res, bkh_info *inf) {
    hipStream_t s = t->stream;
    const long long N = B->N;
    const int ns = B->ns;
    constexpr int NG = (L + 1) * (L + 2) / 2;
    const size_t nparts = (size_t)std::max(std::max((long long)B->nunit, (long long)B->nitem) * NG,
                                           (long long)B->ntile * 2);
    const int K = o.graph_iters > 0 ? o.graph_iters : 4;
    const bool ir = o.precision == BKH_FP32_IR;
    if (ir && !B->val32) { bkh_set_error("FP32_IR needs a batch created with want_fp32"); return BKH_BAD_ARGUMENT; }
    const size_t vbytes = ir ? sizeof(float) : sizeof(double);
    // path 0 / 1: one launch per phase in hipGraphs; 2: one persistent launch per solve.
    // 512 uneven systems (14.8M rows, 251M nonzeros), BiCGSTAB(2) to 1e-8, chunked SpMV:
    // graphs 95.2 ms FP64 and 64.6 ms FP32-IR, persistent 112.4 and 72.0 ms (CSR-stream
    // SpMV: graphs 173.9 / 148.1, persistent FP32-IR 192.0); x is bitwise the same on
    // both paths with the same SpMV.
    const bool persist = o.path == 2;
    size_t bytes = carve_size(sizeof(Ctl)) + carve_size(sizeof(BtGrid)) + carve_size(sizeof(BtSys) * ns) + carve_size(sizeof(double) * nparts) +
                   carve_size(sizeof(double) * ns) + carve_size(sizeof(int) * ns) + carve_size(sizeof(double) * ns) +
                   carve_size(sizeof(double) * N) + (2 * (L + 1) + 3) * carve_size(vbytes * N);
    void *ws = bkh_workspace(bytes);
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1);
    BtGrid *grid = cv.take<BtGrid>(1);
    BtSys *sy = cv.take<BtSys>(ns);
    double *part = cv.take<double>(nparts);
    double *tol_d = cv.take<double>(ns);
    int *frozen_d = cv.take<int>(ns);
    double *scale_d = cv.take<double>(ns);
    double *r64 = cv.take<double>(N);
    std::vector<BtSys> hs(ns);
    std::vector<int> total_it(ns, 0);
    HIP_OK(hipEventRecord(t->ev0, s));
    int rc = 0, passes = 0;
    if (!ir) {
        double *R = cv.take<double>((L + 1) * N), *P = cv.take<double>((L + 1) * N), *rs = cv.take<double>(N);
        hipLaunchKernelGGL(k_bt_sys_init, dim3((ns + 255) / 256), dim3(256), 0, s, sy, ns, (const double *)nullptr,
                           tol, (const int *)nullptr, max_iter);
        rc = bt_core<L, double, double>(t, B, B->val, b, x, R, P, rs, sy, c, part, max_iter, K,
                                        persist ? grid : nullptr);
        if (rc) return rc;
        HIP_OK(hipMemcpyAsync(hs.data(), sy, sizeof(BtSys) * ns, hipMemcpyDeviceToHost, s)); HIP_OK(hipStreamSynchronize(s));
        for (int q = 0; q < ns; q++) {
            ret[q] = hs[q].status == ST_CONVERGED ? hs[q].result : (hs[q].status == ST_BREAKDOWN ? -2 : -1);
            total_it[q] = hs[q].it;
        }
        passes = 1;
    } else {
        float *R = cv.take<float>((L + 1) * N), *P = cv.take<float>((L + 1) * N), *rs = cv.take<float>(N);
        float *r32 = cv.take<float>(N), *d32 = cv.take<float>(N);
        // each pass's inner target: at least 1e-4 relative (512 uneven systems, tol 1e-8:
        // 1e-4 -> 2 passes / 29 inner iterations / 61.6 ms; 1e-5 -> 31 / 64.4 ms;
        // 1e-6 -> 34 / 69.2 ms: a restarted FP32 Krylov space costs fewer iterations than
        // one pushed towards FP32's floor); the FP64 true residual decides every stop
        const double itol_min = o.inner_tol > 0.0 ? o.inner_tol : 1e-4;
        const int max_pass = o.max_inner > 0 ? o.max_inner : 12;
        std::vector<double> tol_h(ns), scale_h(ns);
        std::vector<int> frozen_h(ns), done_h(ns, 0);
        for (;;) {
            // FP64 true residual of every system
            hipLaunchKernelGGL(k_bt_sys_init, dim3((ns + 255) / 256), dim3(256), 0, s, sy, ns,
                               (const double *)nullptr, tol, (const int *)nullptr, 1);
            if ((rc = bt_true_res(t, B, b, x, r64, sy, part))) return rc;
            HIP_OK(hipMemcpyAsync(hs.data(), sy, sizeof(BtSys) * ns, hipMemcpyDeviceToHost, s)); HIP_OK(hipStreamSynchronize(s));
            int active = 0;
            for (int q = 0; q < ns; q++) {
                const double rel = hs[q].rel;
                const bool conv = rel < tol;
                const bool spent = total_it[q] >= max_iter || done_h[q] < 0;
                frozen_h[q] = (conv || spent) ? 1 : 0;
                if (conv) ret[q] = total_it[q] > 0 ? total_it[q] - 1 : 0;
                else if (spent) ret[q] = (done_h[q] == -2) ? -2 : -1;
                if (true_res) true_res[q] = rel;
                // inner system scaled to unit norm; relative inner target
                scale_h[q] = hs[q].err > 0.0 ? 1.0 / hs[q].err : 1.0;
                tol_h[q] = std::max(itol_min, 0.5 * tol / std::max(rel, 1e-300));
                active += frozen_h[q] ? 0 : 1;
            }
            if (active == 0 || passes >= max_pass) {
                for (int q = 0; q < ns; q++)
                    if (!frozen_h[q]) ret[q] = -1;
                break;
            }
            HIP_OK(hipMemcpyAsync(tol_d, tol_h.data(), sizeof(double) * ns, hipMemcpyHostToDevice, s)); HIP_OK(hipStreamSynchronize(s));
            HIP_OK(hipMemcpyAsync(frozen_d, frozen_h.data(), sizeof(int) * ns, hipMemcpyHostToDevice, s)); HIP_OK(hipStreamSynchronize(s));
            HIP_OK(hipMemcpyAsync(scale_d, scale_h.data(), sizeof(double) * ns, hipMemcpyHostToDevice, s)); HIP_OK(hipStreamSynchronize(s));
            hipLaunchKernelGGL(k_bt_ir_prep, dim3(B->ntile), dim3(256), 0, s, B->trow, B->tsys, (const double *)r64,
                               r32, d32, (const double *)scale_d);
            hipLaunchKernelGGL(k_bt_sys_init, dim3((ns + 255) / 256), dim3(256), 0, s, sy, ns, (const double *)tol_d,
                               0.0, (const int *)frozen_d, max_iter);
            rc = bt_core<L, float, float>(t, B, B->val32, r32, d32, R, P, rs, sy, c, part, max_iter, K,
                                          persist ? grid : nullptr);
            if (rc) return rc;
            HIP_OK(hipMemcpyAsync(hs.data(), sy, sizeof(BtSys) * ns, hipMemcpyDeviceToHost, s)); HIP_OK(hipStreamSynchronize(s));
            for (int q = 0; q < ns; q++) {
                if (frozen_h[q]) continue;
                total_it[q] += hs[q].it;
                if (hs[q].status == ST_BREAKDOWN && hs[q].it == 0) done_h[q] = -2;
            }
            hipLaunchKernelGGL(k_bt_ir_update, dim3(B->ntile), dim3(256), 0, s, B->trow, B->tsys, x,
                               (const float *)d32, (const double *)scale_d);
            LAUNCH_OK();
            passes++;
        }
    }
    HIP_OK(hipEventRecord(t->ev1, s));
    HIP_OK(hipEventSynchronize(t->ev1));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    bkh_note_solve_ms(ms);
    if (!ir && true_res) {
        if ((rc = bt_true_res(t, B, b, x, r64, sy, part))) return rc;
        HIP_OK(hipMemcpyAsync(hs.data(), sy, sizeof(BtSys) * ns, hipMemcpyDeviceToHost, s)); HIP_OK(hipStreamSynchronize(s));
        for (int q = 0; q < ns; q++) true_res[q] = hs[q].rel;
    }
    int worst = 0, conv = 0, maxit_ = 0;
    for (int q = 0; q < ns; q++) {
        if (ret[q] >= 0) { conv++; maxit_ = std::max(maxit_, ret[q]); }
        else worst = std::min(worst, ret[q]);
    }
    if (inf) {
        memset(inf, 0, sizeof *inf);
        inf->time_ms = ms;
        inf->iters = maxit_;
        inf->passes = passes;
        int mx = 0;
        for (int q = 0; q < ns; q++) mx = std::max(mx, total_it[q]);
        inf->inner_iters = mx;
        inf->status = conv == ns ? ST_CONVERGED : (worst == -2 ? ST_BREAKDOWN : ST_MAXIT);
        inf->ret = conv == ns ? maxit_ : worst;
        if (true_res) {
            double mxr = 0.0;
            for (int q = 0; q < ns; q++) mxr = std::max(mxr, true_res[q]);
            inf->true_res = mxr;
        }
    }
    return conv == ns ? maxit_ : worst;
}

extern "C" int bkh_batch_bicgstabl(bkh_batch *B, const double *d_b, int l, double *d_x, double tol, int max_iter,
                                   const bkh_opts *opts, int *ret, double *true_res, bkh_info *info) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!B || !d_b || !d_x || !ret) return BKH_BAD_ARGUMENT;
    if (l < 1 || l > 4) { bkh_set_error("batched BiCGSTAB(l) supports 1 <= l <= 4"); return -3; }
    bkh_opts o = opts ? *opts : bkh_opts{};
    switch (l) {
    case 1: return bt_solve<1>(t, B, d_b, d_x, tol, max_iter, o, ret, true_res, info);
    case 2: return bt_solve<2>(t, B, d_b, d_x, tol, max_iter, o, ret, true_res, info);
    case 3: return bt_solve<3>(t, B, d_b, d_x, tol, max_iter, o, ret, true_res, info);
    default: return bt_solve<4>(t, B, d_b, d_x, tol, max_iter, o, ret, true_res, info);
    }
}

extern "C" int bkh_batch_last_path(void) { return g_bt_last_path; }

// y = A x over the whole batch (FP64, or FP32 values and vectors).
extern "C" int bkh_batch_spmv(const bkh_batch *B, int fp32, const void *x, void *y) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!B || !x || !y || (fp32 && !B->val32)) return BKH_BAD_ARGUMENT;
    hipStream_t s = t->stream;
    if (fp32) {
        Epi<0, float> e;
        memset(&e, 0, sizeof e);
        bt_spmv<float, float, 0>(s, B, B->val32, nullptr, (const float *)x, (float *)y, e, nullptr);
    } else {
        Epi<0, double> e;
        memset(&e, 0, sizeof e);
        bt_spmv<double, double, 0>(s, B, B->val, nullptr, (const double *)x, (double *)y, e, nullptr);
    }
    LAUNCH_OK();
    HIP_OK(hipStreamSynchronize(s));
    return 0;
}

// Autotuner: times the batched SpMV for candidate blockings (rows per wavefront,
// nonzeros per wavefront, the latter derived from the batch's row-length histogram)
// and a vector kernel for candidate tile sizes, on this batch, and keeps the fastest
// layout. fp32 selects the FP32 SpMV (the inner
// iteration of BKH_FP32_IR). best_ms receives the chosen SpMV's time.
extern "C" int bkh_batch_tune(bkh_batch *B, int fp32, double *best_ms) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!B || (fp32 && !B->val32)) return BKH_BAD_ARGUMENT;
    hipStream_t s = t->stream;
    const long long N = B->N;
    const size_t vb = fp32 ? sizeof(float) : sizeof(double);
    void *ws = bkh_workspace(4 * carve_size(vb * N) + carve_size(sizeof(BtSys) * B->ns));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    char *x = cv.take<char>(vb * N), *y = cv.take<char>(vb * N), *z = cv.take<char>(2 * vb * N);
    BtSys *sy = cv.take<BtSys>(B->ns);
    HIP_OK(hipMemsetAsync(x, 0, vb * N, s));
    HIP_OK(hipMemsetAsync(z, 0, 2 * vb * N, s));
    HIP_OK(hipMemsetAsync(sy, 0, sizeof(BtSys) * B->ns, s));
    // SpMV blockings from the batch's row-length histogram: for rows per wavefront
    // 16, 32 and 64, nonzero caps at that many rows of median, 90th-percentile and
    // mean length (rounded up to 64, at least the longest row, at most SPMV_WCAP),
    // and SPMV_WCAP itself
    std::vector<int> lens(B->N);
    long long tot = 0;
    for (long long i = 0; i < N; i++) { lens[i] = B->rp_h[i + 1] - B->rp_h[i]; tot += lens[i]; }
    std::vector<int> sorted_lens = lens;
    std::sort(sorted_lens.begin(), sorted_lens.end());
    const double p50 = N ? sorted_lens[N / 2] : 1, p90 = N ? sorted_lens[(N * 9) / 10] : 1;
    const double mean = N ? (double)tot / N : 1;
    std::vector<std::array<int, 2>> cand;
    for (int rc : {16, 32, 64})
        for (double per : {p50, p90, mean, (double)SPMV_WCAP / rc}) {
            int cap = (int)std::ceil(rc * per / 64.0) * 64;
            cap = std::min(SPMV_WCAP, std::max(cap, std::max(B->max_row, 64)));
            const std::array<int, 2> c = {rc, cap};
            if (std::find(cand.begin(), cand.end(), c) == cand.end()) cand.push_back(c);
        }
    const int tiles[] = {512, 1024, 2048, 4096};
    const int reps = 5;
    double best = 1e300;
    int bi = 0;
    auto time_spmv = [&]() -> double {
        for (int r = 0; r < reps + 1; r++) {
            if (r == 1) hipEventRecord(t->ev0, s);
            if (fp32) {
                Epi<0, float> e;
                memset(&e, 0, sizeof e);
                bt_spmv<float, float, 0>(s, B, B->val32, nullptr, (const float *)x, (float *)y, e, nullptr);
            } else {
                Epi<0, double> e;
                memset(&e, 0, sizeof e);
                bt_spmv<double, double, 0>(s, B, B->val, nullptr, (const double *)x, (double *)y, e, nullptr);
            }
        }
        hipEventRecord(t->ev1, s);
        hipEventSynchronize(t->ev1);
        float ms = 0.f;
        hipEventElapsedTime(&ms, t->ev0, t->ev1);
        return ms / reps;
    };
    B->use_lc = 0;
    for (int q = 0; q < (int)cand.size(); q++) {
        if (bt_layout(B, cand[q][0], cand[q][1], B->tile_rows)) return BKH_DEVICE_ERROR;
        const double ms = time_spmv();
        if (ms < best) { best = ms; bi = q; }
    }
    if (bt_layout(B, cand[bi][0], cand[bi][1], B->tile_rows)) return BKH_DEVICE_ERROR;
    if (B->lc_ok) {                                       // the column-chunked kernel against the best blocking
        B->use_lc = 1;
        const double ms = time_spmv();
        if (ms < best) best = ms;
        else B->use_lc = 0;
    }
    bkh_graph_epoch_bump();
    if (best_ms) *best_ms = best;
    // vector tiles: time the BiCG update kernel (r -= alpha p, x += alpha p) per size
    double bt = 1e300;
    int ti = 1;
    for (int q = 0; q < (int)(sizeof tiles / sizeof tiles[0]); q++) {
        if (bt_layout(B, cand[bi][0], cand[bi][1], tiles[q])) return BKH_DEVICE_ERROR;
        for (int r = 0; r < reps + 1; r++) {
            if (r == 1) hipEventRecord(t->ev0, s);
            if (fp32)
                hipLaunchKernelGGL(k_bt_r<float>, dim3(B->ntile), dim3(256), 0, s, B->trow, B->tsys, N, 0,
                                   (const float *)z, (float *)y, (float *)x, (const BtSys *)sy);
            else
                hipLaunchKernelGGL(k_bt_r<double>, dim3(B->ntile), dim3(256), 0, s, B->trow, B->tsys, N, 0,
                                   (const double *)z, (double *)y, (double *)x, (const BtSys *)sy);
        }
        hipEventRecord(t->ev1, s);
        hipEventSynchronize(t->ev1);
        float ms = 0.f;
        hipEventElapsedTime(&ms, t->ev0, t->ev1);
        if (ms < bt) { bt = ms; ti = q; }
    }
    if (bt_layout(B, cand[bi][0], cand[bi][1], tiles[ti])) return BKH_DEVICE_ERROR;
    LAUNCH_OK();
    return 0;
}
