// bkh_spmv.hpp -- CSR SpMV kernel templates (gfx950).
//
// CSR-stream layout: each workgroup owns a contiguous row range whose nonzeros
// fit in LDS. Phase 1 streams val/col with coalesced 16-byte loads (plain loads:
// the non-temporal hint measured ~12% slower on MI355X), gathers
// x (L2/MALL-resident for banded matrices) and stores the products in LDS.
// Phase 2 gives each thread one row and adds that row's products in stored
// order, so y[i] = (((0 + a_0 x_0) + a_1 x_1) + ...) exactly as src/matvec.c:22-25
// computes it. A row longer than the LDS capacity gets a workgroup to itself and
// a tree reduction. An optional epilogue accumulates NV dot products between
// y and up to 8 vectors read at the same rows (the Krylov scalars that follow
// an SpMV), reduced deterministically with a solver-specific finalizer.
#pragma once
#include "bkh_core.hpp"
#include "bkh_host.hpp"
#include <initializer_list>
#include <utility>

#define EPI_Y 15   // epilogue operand index meaning "the freshly computed y"

template <int NV, typename WT = double>
struct Epi {
    const WT *w[8];
    unsigned char a[NV > 0 ? NV : 1], b[NV > 0 ? NV : 1];
};

struct NoFin {
    __device__ void operator()(Ctl *, const double *) const {}
};

template <typename VT, typename YT>
struct SpmvArgs {
    const int *rp, *ci, *blk;
    const VT *val;
    const YT *x;
    YT *y;
    double alpha, beta;
    int alpha_one, has_beta;
};

template <int NV, typename WT>
__device__ __forceinline__ double epi_operand(const Epi<NV, WT> &e, int idx, int row, double y) {
    return idx == EPI_Y ? y : (double)e.w[idx][row];
}

template <typename VT, typename YT, int NV, class Fin>
__global__ __launch_bounds__(SPMV_BS) void k_spmv(SpmvArgs<VT, YT> a, Epi<NV, YT> e, Ctl *ctl,
                                                  double *part, Fin fin) {
    if (ctl && ctl_stopped(ctl)) return;
    __shared__ YT prod[SPMV_CAP];
    const int tid = threadIdx.x;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = a.blk[b], r1 = a.blk[b + 1];
    const int k0 = a.rp[r0], k1 = a.rp[r1];
    double dv[NV > 0 ? NV : 1];
#pragma unroll
    for (int i = 0; i < NV; i++) dv[i] = 0.0;

    if (k1 - k0 <= SPMV_CAP) {
        // 16-byte loads of values (VEC per thread per step) and matching column
        // indices, from an aligned base; every step is predicated so the whole
        // stage unrolls and all loads are in flight before the x gathers.
        constexpr int VEC = 16 / sizeof(VT);
        constexpr int STEPS = SPMV_CAP / (VEC * SPMV_BS) + 1;
        typedef VT vvec __attribute__((ext_vector_type(VEC)));
        typedef int ivec __attribute__((ext_vector_type(VEC)));
        const YT al = (YT)a.alpha;
        const int kb = k0 & ~(VEC - 1);
        vvec vv[STEPS];
        ivec cc[STEPS];
#pragma unroll
        for (int j = 0; j < STEPS; j++) {
            const int k = kb + VEC * (tid + j * SPMV_BS);
            if (k < k1) {
                vv[j] = *(const vvec *)(a.val + k);
                cc[j] = *(const ivec *)(a.ci + k);
            }
        }
#pragma unroll
        for (int j = 0; j < STEPS; j++) {
            const int k = kb + VEC * (tid + j * SPMV_BS);
#pragma unroll
            for (int q = 0; q < VEC; q++) {
                if (k + q >= k0 && k + q < k1) {
                    YT v = (YT)vv[j][q];
                    YT xv = a.x[cc[j][q]];
                    prod[k + q - k0] = a.alpha_one ? mul_rn(v, xv) : mul_rn(mul_rn(al, v), xv);
                }
            }
        }
        __syncthreads();
        const int row = r0 + tid;
        if (row < r1) {
            const int s0 = a.rp[row] - k0, s1 = a.rp[row + 1] - k0;
            YT s = a.has_beta ? mul_rn((YT)a.beta, a.y[row]) : (YT)0;
            for (int k = s0; k < s1; k++) s = add_rn(s, prod[k]);
            a.y[row] = s;
#pragma unroll
            for (int i = 0; i < NV; i++)
                dv[i] = epi_operand(e, e.a[i], row, (double)s) * epi_operand(e, e.b[i], row, (double)s);
        }
    } else {
        // one long row: strided partial sums + deterministic tree reduction
        double acc[1] = {0.0};
        for (int k = k0 + tid; k < k1; k += SPMV_BS) {
            double v = (double)a.val[k];
            if (!a.alpha_one) v = a.alpha * v;
            acc[0] += v * (double)a.x[a.ci[k]];
        }
        __shared__ double shl[SPMV_BS / BKH_WAVE];
        block_sum<1, SPMV_BS>(acc, shl);
        __shared__ double ybc;
        if (tid == 0) {
            double s = acc[0];
            if (a.has_beta) s += a.beta * (double)a.y[r0];
            a.y[r0] = (YT)s;
            ybc = (double)(YT)s;
        }
        __syncthreads();
        if (tid == 0) {
#pragma unroll
            for (int i = 0; i < NV; i++)
                dv[i] = epi_operand(e, e.a[i], r0, ybc) * epi_operand(e, e.b[i], r0, ybc);
        }
    }
    if constexpr (NV > 0) block_partial<NV, SPMV_BS>(dv, part);
}

// Wave-stream SpMV: as k_spmv, but every wave streams its own row block into
// its own LDS slice, so the four waves of a workgroup never wait on each other.
template <typename VT, typename YT, int NV, class Fin>
__global__ __launch_bounds__(256) void k_spmv_wave(SpmvArgs<VT, YT> a, int nwblk, Epi<NV, YT> e,
                                                   Ctl *ctl, double *part, Fin fin) {
    if (ctl && ctl_stopped(ctl)) return;
    __shared__ YT prod_all[4][SPMV_WCAP];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    YT *prod = prod_all[w];
    const int b = xcd_block(blockIdx.x, gridDim.x) * 4 + w;
    double dv[NV > 0 ? NV : 1];
#pragma unroll
    for (int i = 0; i < NV; i++) dv[i] = 0.0;
    if (b < nwblk) {
        const int r0 = a.blk[b], r1 = a.blk[b + 1];
        const int k0 = a.rp[r0], k1 = a.rp[r1];
        if (k1 - k0 <= SPMV_WCAP) {
            constexpr int VEC = 16 / sizeof(VT);
            constexpr int STEPS = SPMV_WCAP / (VEC * 64) + 1;
            typedef VT vvec __attribute__((ext_vector_type(VEC)));
            typedef int ivec __attribute__((ext_vector_type(VEC)));
            const YT al = (YT)a.alpha;
            const int kb = k0 & ~(VEC - 1);
            vvec vv[STEPS];
            ivec cc[STEPS];
#pragma unroll
            for (int j = 0; j < STEPS; j++) {
                const int k = kb + VEC * (lane + j * 64);
                if (k < k1) {
                    vv[j] = *(const vvec *)(a.val + k);
                    cc[j] = *(const ivec *)(a.ci + k);
                }
            }
#pragma unroll
            for (int j = 0; j < STEPS; j++) {
                const int k = kb + VEC * (lane + j * 64);
#pragma unroll
                for (int q = 0; q < VEC; q++) {
                    if (k + q >= k0 && k + q < k1) {
                        YT v = (YT)vv[j][q];
                        YT xv = a.x[cc[j][q]];
                        prod[k + q - k0] = a.alpha_one ? mul_rn(v, xv) : mul_rn(mul_rn(al, v), xv);
                    }
                }
            }
            // LDS writes of this wave complete before its reads (wave-local handoff)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int row = r0 + lane;
            if (row < r1) {
                const int s0 = a.rp[row] - k0, s1 = a.rp[row + 1] - k0;
                YT s = a.has_beta ? mul_rn((YT)a.beta, a.y[row]) : (YT)0;
                for (int k = s0; k < s1; k++) s = add_rn(s, prod[k]);
                a.y[row] = s;
#pragma unroll
                for (int i = 0; i < NV; i++)
                    dv[i] = epi_operand(e, e.a[i], row, (double)s) * epi_operand(e, e.b[i], row, (double)s);
            }
        } else {
            // one long row per wave: strided sums, butterfly reduction
            double acc = 0.0;
            for (int k = k0 + lane; k < k1; k += 64) {
                double v = (double)a.val[k];
                if (!a.alpha_one) v = a.alpha * v;
                acc += v * (double)a.x[a.ci[k]];
            }
            acc = wave_sum(acc);
            if (a.has_beta) acc += a.beta * (double)a.y[r0];
            const double ys = (double)(YT)acc;
            if (lane == 0) {
                a.y[r0] = (YT)acc;
#pragma unroll
                for (int i = 0; i < NV; i++)
                    dv[i] = epi_operand(e, e.a[i], r0, ys) * epi_operand(e, e.b[i], r0, ys);
            }
        }
    }
    if constexpr (NV > 0) block_partial<NV, 256>(dv, part);
}

// SELL-64 SpMV: one lane per row, four slices (256 rows) per workgroup. Step k of
// a wave loads 64 consecutive values and columns (entry k of each of its rows) and
// gathers x at those columns; each lane adds its row's products in stored order, so
// the result is the sequential row sum of src/matvec.c:22-25 bit for bit. IT is int
// (columns) or short (column - row). DIA: the slice-diagonal copy (so, val, col = its
// slice offsets, values and per-slice column offsets, uniform over the wave's slice);
// its value-0 positions add 0 products at clamped columns, which leaves every sum as on
// the SELL copy for finite x (the solvers' own products).
template <typename VT, typename YT, typename IT, int NV, class Fin, bool DIA = false>
__global__ __launch_bounds__(256) void k_spmv_sell(const int *__restrict__ rp, const int *__restrict__ so,
                                                   const VT *__restrict__ val, const IT *__restrict__ col, int n,
                                                   const YT *x, YT *y, double alpha, double beta, int alpha_one,
                                                   int has_beta, Epi<NV, YT> e, Ctl *ctl, double *part, Fin fin) {
    if (ctl && ctl_stopped(ctl)) return;
    const int row = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    double dv[NV > 0 ? NV : 1];
#pragma unroll
    for (int i = 0; i < NV; i++) dv[i] = 0.0;
    if (row < n) {
        int len, base;
        const int *dof = nullptr;
        if constexpr (DIA) {
            const int sl = __builtin_amdgcn_readfirstlane(row >> 6);
            const int s0 = so[sl];
            len = (so[sl + 1] - s0) >> 6;
            base = s0 + (row & 63);
            dof = (const int *)col + (s0 >> 6);
        } else {
            len = rp[row + 1] - rp[row];
            base = so[row >> 6] + (row & 63);
        }
        const YT al = (YT)alpha;
        YT s = has_beta ? mul_rn((YT)beta, y[row]) : (YT)0;
#pragma unroll 4
        for (int k = 0; k < len; k++) {
            const int ei = base + 64 * k;
            int c;
            if constexpr (DIA) c = min(max(row + dof[k], 0), n - 1);
            else c = sizeof(IT) == 2 ? row + (int)col[ei] : (int)col[ei];
            const YT v = (YT)val[ei];
            s = add_rn(s, alpha_one ? mul_rn(v, x[c]) : mul_rn(mul_rn(al, v), x[c]));
        }
        y[row] = s;
#pragma unroll
        for (int i = 0; i < NV; i++)
            dv[i] = epi_operand(e, e.a[i], row, (double)s) * epi_operand(e, e.b[i], row, (double)s);
    }
    if constexpr (NV > 0) block_partial<NV, 256>(dv, part);
}

// Number of workgroups a launch uses (also the number of partials it writes).
static inline int spmv_grid(const bkh_matrix *A) {
    if (A->variant == SPMV_SELL || A->variant == SPMV_SELL32) return (A->n + 255) / 256;
    return A->variant == SPMV_WAVE ? (A->nwblk + 3) / 4 : A->nblk;
}

// Launches the SpMV kernel only (partials written, no finalize); returns the
// number of workgroups, i.e. of partial rows. dia: the solvers' own products (finite
// x) run on the slice-diagonal copy when the matrix has one and the default SELL-64
// kernel is selected; the public SpMV entry points pass false.
template <typename VT, typename YT, int NV, class Fin>
static inline int launch_spmv_kernel(hipStream_t s, const bkh_matrix *A, const VT *val, const YT *x,
                                     YT *y, double alpha, double beta, const Epi<NV, YT> &e, Ctl *ctl,
                                     double *part, Fin fin, bool dia = true) {
    SpmvArgs<VT, YT> a;
    a.rp = A->rp; a.ci = A->ci; a.blk = A->blk; a.val = val; a.x = x; a.y = y;
    a.alpha = alpha; a.beta = beta; a.alpha_one = (alpha == 1.0); a.has_beta = (beta != 0.0);
    const int G = spmv_grid(A);
    if (dia && A->dia_entries > 0 && A->nsl > 0 && A->variant == SPMV_SELL) {
        const VT *dvv;
        if constexpr (sizeof(VT) == sizeof(float)) dvv = (const VT *)A->dval32;
        else dvv = (const VT *)A->dval;
        if (dvv) {
            hipLaunchKernelGGL((k_spmv_sell<VT, YT, int, NV, Fin, true>), dim3(G), dim3(256), 0, s, A->rp, A->dso,
                               dvv, (const int *)A->doff, A->n, x, y, alpha, beta, (int)(alpha == 1.0),
                               (int)(beta != 0.0), e, ctl, part, fin);
            return G;
        }
    }
    if (A->variant == SPMV_SELL || A->variant == SPMV_SELL32) {
        const VT *sv;
        if constexpr (sizeof(VT) == sizeof(float)) sv = (const VT *)A->sval32;
        else sv = (const VT *)A->sval;
        if (A->variant == SPMV_SELL && A->scd)
            hipLaunchKernelGGL((k_spmv_sell<VT, YT, short, NV, Fin>), dim3(G), dim3(256), 0, s, A->rp, A->sso, sv,
                               (const short *)A->scd, A->n, x, y, alpha, beta, (int)(alpha == 1.0),
                               (int)(beta != 0.0), e, ctl, part, fin);
        else
            hipLaunchKernelGGL((k_spmv_sell<VT, YT, int, NV, Fin>), dim3(G), dim3(256), 0, s, A->rp, A->sso, sv,
                               (const int *)A->scol, A->n, x, y, alpha, beta, (int)(alpha == 1.0),
                               (int)(beta != 0.0), e, ctl, part, fin);
        return G;
    }
    if (A->variant == SPMV_WAVE) {
        a.blk = A->wblk;
        hipLaunchKernelGGL((k_spmv_wave<VT, YT, NV, Fin>), dim3(G), dim3(256), 0, s, a, A->nwblk, e,
                           ctl, part, fin);
    } else {
        hipLaunchKernelGGL((k_spmv<VT, YT, NV, Fin>), dim3(G), dim3(SPMV_BS), 0, s, a, e, ctl,
                           part, fin);
    }
    return G;
}

// Host launcher: SpMV plus, when NV > 0, the finalize that reduces its epilogue
// dots and runs `fin`. `ctl`/`part` may be null when NV == 0.
template <typename VT, typename YT, int NV, class Fin>
static inline void launch_spmv(hipStream_t s, const bkh_matrix *A, const VT *val, const YT *x,
                               YT *y, double alpha, double beta, const Epi<NV, YT> &e, Ctl *ctl,
                               double *part, Fin fin, bool dia = true) {
    const int G = launch_spmv_kernel<VT, YT, NV>(s, A, val, x, y, alpha, beta, e, ctl, part, fin, dia);
    if constexpr (NV > 0) launch_finalize<NV>(s, part, G, ctl, fin);
}

template <int NV, typename WT = double>
static inline Epi<NV, WT> make_epi(std::initializer_list<const WT *> w,
                                   std::initializer_list<std::pair<int, int>> pairs) {
    Epi<NV, WT> e;
    memset(&e, 0, sizeof(e));
    int i = 0;
    for (auto p : w) e.w[i++] = p;
    i = 0;
    for (auto pr : pairs) { e.a[i] = (unsigned char)pr.first; e.b[i] = (unsigned char)pr.second; i++; }
    return e;
}
