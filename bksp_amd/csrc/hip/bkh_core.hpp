// bkh_core.hpp -- internal header of the MI355X (gfx950) device layer.
//
// Device-side building blocks shared by every kernel:
//   * Ctl: the per-solve control block in device memory. Kernels read `done` at
//     entry and return at once when a solve has stopped, so a captured graph of
//     K iterations can be replayed without a host round trip per iteration and
//     the exact break index is still recorded on the device.
//   * deterministic reductions: wave64 shuffles -> LDS -> one partial per
//     workgroup -> a one-workgroup finalize kernel sums the partials in a fixed
//     order and runs the solver's scalar update. No atomics, so repeated runs
//     are bitwise identical.
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
    int base;          // loop index of the first iteration of a graph replay
    int first;         // loop index at which the current pass started
    double tol;
    double fnb;        // norm of the right-hand side
    double err;        // last recursive residual ratio
    double pad1;
    double s[BKH_MAX_SCAL];   // solver scalars (layout owned by each solver)
};

// ---------------------------------------------------------------- reductions

__device__ __forceinline__ bool ctl_stopped(const Ctl *c) {
    return __builtin_nontemporal_load(&c->done) != 0;
}


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

// Writes this workgroup's NV partial sums to part[blockIdx.x * NV + i]. The
// grid totals are formed by k_finalize, launched right after the producing
// kernel: the kernel boundary orders and publishes the partials, so no
// device-scope fence or atomic is needed inside the producer (on a multi-XCD
// part each such fence would write back the XCD's L2).
template <int NV, int BS>
__device__ __forceinline__ void block_partial(double (&v)[NV], double *part) {
    __shared__ double sh[NV * (BS / BKH_WAVE)];
    block_sum<NV, BS>(v, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; i++) part[(size_t)blockIdx.x * NV + i] = v[i];
    }
}

// First level of a two-level partial reduction: workgroup r sums partial rows
// [r*RED_CHUNK, (r+1)*RED_CHUNK) of an nblocks x E array (E <= 256 values per
// row) into out[r*E + e], in a fixed order.
#define RED_CHUNK 128
__global__ __launch_bounds__(256) void k_part_reduce(const double *part, int nblocks, int E, double *out);

// Sums the nblocks rows of NV partials in a fixed order; valid in thread 0 of the
// calling workgroup (256 threads).
template <int NV>
__device__ __forceinline__ void sum_partials(const double *part, int nblocks, double (&acc)[NV]) {
    __shared__ double sh[NV * 4];
#pragma unroll
    for (int i = 0; i < NV; i++) acc[i] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 256) {
#pragma unroll
        for (int i = 0; i < NV; i++) acc[i] += part[(size_t)b * NV + i];
    }
    block_sum<NV, 256>(acc, sh);
}

// One workgroup: sums the partials in a fixed order (deterministic) and runs the
// solver's scalar update `fin` in thread 0.
template <int NV, class Fin>
__global__ __launch_bounds__(256) void k_finalize(const double *part, int nblocks, Ctl *c, Fin fin) {
    if (c && ctl_stopped(c)) return;
    double acc[NV];
    sum_partials<NV>(part, nblocks, acc);
    if (threadIdx.x == 0) fin(c, acc);
}

// Second-level buffer of the thread's stream (see k_part_reduce); nullptr if the
// runtime is not initialised.
double *bkh_red2();

// Launches the finalize of nblocks x NV partials; wide ones go through
// k_part_reduce first so no thread sums more than a few hundred values serially.
template <int NV, class Fin>
static inline void launch_finalize(hipStream_t s, const double *part, int nblocks, Ctl *c, Fin fin) {
    if ((long long)nblocks * NV > 8192) {
        double *r2 = bkh_red2();
        const int nb2 = (nblocks + RED_CHUNK - 1) / RED_CHUNK;
        hipLaunchKernelGGL(k_part_reduce, dim3(nb2), dim3(256), 0, s, part, nblocks, NV, r2);
        hipLaunchKernelGGL((k_finalize<NV, Fin>), dim3(1), dim3(256), 0, s, (const double *)r2, nb2, c, fin);
    } else {
        hipLaunchKernelGGL((k_finalize<NV, Fin>), dim3(1), dim3(256), 0, s, part, nblocks, c, fin);
    }
}

// Logical workgroup index: XCD x (hardware dispatch blockIdx % 8) receives the
// contiguous logical range [start_x, start_x + count_x).
__device__ __forceinline__ int xcd_block(int b, int G) {
    const int x = b % BKH_NXCD, i = b / BKH_NXCD;
    const int per = G / BKH_NXCD, rem = G % BKH_NXCD;
    return (x < rem) ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
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
