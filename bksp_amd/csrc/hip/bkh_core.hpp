// bkh_core.hpp -- internal header of the MI355X (gfx950) device layer.
//
// Device-side building blocks shared by every kernel:
//   * Ctl: the per-solve control block in device memory. Kernels read `done` at
//     entry and return at once when a solve has stopped, so a captured graph of
//     K iterations can be replayed without a host round trip per iteration and
//     the exact break index is still recorded on the device.
//   * deterministic reductions: wave64 shuffles -> LDS -> one partial per
//     workgroup -> the last workgroup to finish (atomic ticket) sums the partials
//     in a fixed order and runs the solver's scalar update. No float atomics, so
//     repeated runs are bitwise identical.
//   * xcd_block(): maps the hardware's round-robin workgroup-to-XCD dispatch
//     onto contiguous row ranges per XCD, so each XCD's private L2 sees
//     neighbouring rows (and their x entries) together.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#define BKH_WAVE 64
#define BKH_NXCD 8
#define BKH_MAX_SCAL 160

// Status values stored in Ctl::status.
enum { ST_RUNNING = 0, ST_CONVERGED = 1, ST_MAXIT = 2, ST_BREAKDOWN = 3 };

struct alignas(64) Ctl {
    int done;          // 1: stop (converged, max_it reached or breakdown)
    int status;        // ST_*
    int it;            // loop index of the iteration being executed
    int result_it;     // loop index at the break
    int max_it;
    int inner_total;   // iterations summed over refinement passes
    unsigned ticket;   // grid-reduction ticket (reset by the last workgroup)
    int pad0;
    double tol;
    double fnb;        // norm of the right-hand side
    double err;        // last recursive residual ratio
    double pad1;
    double s[BKH_MAX_SCAL];   // solver scalars (layout owned by each solver)
};

// ---------------------------------------------------------------- reductions

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sums NV values over the workgroup; the totals are valid in thread 0.
// `sh` must hold NV * (BS/64) doubles.
template <int NV, int BS>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *sh) {
    constexpr int NW = BS / BKH_WAVE;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++) v[i] = wave_sum(v[i]);
    if (NW > 1) {
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i < NV; i++) sh[i * NW + w] = v[i];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
#pragma unroll
            for (int i = 0; i < NV; i++) {
                double t = sh[i * NW];
                for (int k = 1; k < NW; k++) t += sh[i * NW + k];
                v[i] = t;
            }
        }
        __syncthreads();
    }
}

// Grid reduction with a last-workgroup finalizer. Every workgroup of the launch
// must call this exactly once (with its partial values in v). The finalizer runs
// once, in thread 0 of the last workgroup, with the grid totals in `tot`.
// `part` holds gridDim.x * NV doubles.
template <int NV, int BS, class Fin>
__device__ __forceinline__ void grid_sum(double (&v)[NV], double *part, unsigned *ticket,
                                         Fin fin) {
    __shared__ double sh[NV * (BS / BKH_WAVE)];
    __shared__ int is_last;
    block_sum<NV, BS>(v, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; i++) part[(size_t)blockIdx.x * NV + i] = v[i];
        __threadfence();
        unsigned t = atomicAdd(ticket, 1u);
        is_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) acc[i] = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += BS) {
#pragma unroll
        for (int i = 0; i < NV; i++) acc[i] += __builtin_nontemporal_load(&part[(size_t)b * NV + i]);
    }
    block_sum<NV, BS>(acc, sh);
    if (threadIdx.x == 0) {
        fin(acc);
        *ticket = 0u;
        __threadfence();
    }
}

// Logical workgroup index: XCD x (hardware dispatch blockIdx % 8) receives the
// contiguous logical range [start_x, start_x + count_x).
__device__ __forceinline__ int xcd_block(int b, int G) {
    const int x = b % BKH_NXCD, i = b / BKH_NXCD;
    const int per = G / BKH_NXCD, rem = G % BKH_NXCD;
    return (x < rem) ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}

__device__ __forceinline__ bool ctl_stopped(const Ctl *c) {
    return __builtin_nontemporal_load(&c->done) != 0;
}

// Records the end of one iteration: err = sqrt(rr)/fnb; converged if err < tol
// (the reference's test, e.g. cg.c:56-59). NaN/Inf stops with ST_BREAKDOWN.
__device__ __forceinline__ void ctl_check(Ctl *c, double rr) {
    double err = sqrt(rr) / c->fnb;
    c->err = err;
    int it = c->it;
    if (err < c->tol) {
        c->done = 1; c->status = ST_CONVERGED; c->result_it = it;
    } else if (!(err == err) || err > 1.0e300) {
        c->done = 1; c->status = ST_BREAKDOWN; c->result_it = it;
    } else if (it + 1 >= c->max_it) {
        c->done = 1; c->status = ST_MAXIT; c->result_it = it;
    }
    c->it = it + 1;
}

__device__ __forceinline__ void ctl_breakdown(Ctl *c) {
    c->done = 1; c->status = ST_BREAKDOWN; c->result_it = c->it;
}

__device__ __forceinline__ bool bad(double v) { return !(v == v) || v == 0.0 || fabs(v) > 1.0e300; }

// Non-contracted multiply/add: the CSR kernels reproduce the reference's
// separate multiply and sequential add (src/matvec.c:23-24) bit for bit.
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ double add_rn(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ float mul_rn(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ float add_rn(float a, float b) { return __fadd_rn(a, b); }
