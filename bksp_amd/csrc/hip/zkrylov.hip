// zkrylov.hip -- the complex FP64 family on the device:
//   complex CSR products  y = A x, y = A^H x      (replace src/matvec.c:55-98, 174-217)
//   BiCG                                          (replaces src/bicg.c:91-177)
//   BiCGSTAB(l)                                   (replaces src/bicgstabl.c:141-276)
//   shifted BiCGSTAB(l), complex shifts           (replaces src/shbicgstabl.c:288-592)
//
// Complex values are double2 (re, im), the layout of C99 double _Complex, so host
// arrays of the reference's zspmat / double _Complex vectors upload as they are.
// Products follow the reference's C99 arithmetic for finite operands (separate
// multiply and add: re = ac - bd, im = ad + bc); every row of A x and A^H x sums its
// entries in the reference's order (A^H through a transposed copy whose rows list
// the scatter's contributions in source-row order), so the products are bitwise
// equal to src/matvec.c. Scalar divisions use Smith's algorithm.
//
// The solvers follow the reference's BLAS call sequence with the vector operations
// of one step fused into one kernel, every scalar kept on the device (Ctl::s and a
// per-shift array) and one outer iteration captured as a hipGraph. Dot products are
// deterministic (fixed-order partials per workgroup, summed by one workgroup);
// dznrm2 is the square root of that sum of |.|^2.
//
// The reference's zbicgstabl passes a real `one` where cblas_zaxpy reads a complex
// alpha (bicgstabl.c:147, :177, :194); this build uses the intended 1 + 0i.
#include <math.h>
#include <vector>
#include "bkh_host.hpp"
#include "bkh_solver_common.hpp"

typedef double2 zd;

__host__ __device__ __forceinline__ zd zmk(double r, double i) { return make_double2(r, i); }
__device__ __forceinline__ zd zmul(zd a, zd b) {
    const double ac = a.x * b.x, bd = a.y * b.y, ad = a.x * b.y, bc = a.y * b.x;
    return zmk(ac - bd, ad + bc);
}
__device__ __forceinline__ zd zadd(zd a, zd b) { return zmk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ zd zsub(zd a, zd b) { return zmk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ zd zneg(zd a) { return zmk(-a.x, -a.y); }
__device__ __forceinline__ zd zconj(zd a) { return zmk(a.x, -a.y); }
// a / b by Smith's algorithm (scaled by the larger component of b).
__device__ __forceinline__ zd zdiv(zd a, zd b) {
    if (fabs(b.x) >= fabs(b.y)) {
        const double r = b.y / b.x, d = b.x + b.y * r;
        return zmk((a.x + a.y * r) / d, (a.y - a.x * r) / d);
    }
    const double r = b.x / b.y, d = b.x * r + b.y;
    return zmk((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}
__device__ __forceinline__ bool zbad(zd a) {
    const double m = fabs(a.x) + fabs(a.y);
    return !(m == m) || m == 0.0 || m > 1.0e300;
}
__device__ __forceinline__ zd zload(const double *s, int i) { return zmk(s[2 * i], s[2 * i + 1]); }
__device__ __forceinline__ void zstore(double *s, int i, zd v) { s[2 * i] = v.x; s[2 * i + 1] = v.y; }

#define ZLOOP(i, n) \
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < (n); i += (long long)gridDim.x * 256)

// ------------------------------------------------------------------ matrix

// Row-sliced copy of a CSR matrix: slices of 64 rows, entry k of the slice's row l
// at so[slice] + 64 k + l. A slice is as wide as its longest row, but at most
// max(ZS_MIN_W, twice its mean row length), so padding stays within 2 nnz + 64 ZS_MIN_W
// per slice; rl[row] = the row's entries held there (min(length, width)). A row longer
// than its slice's width keeps the rest in a CSR tail: row trow[t] continues with
// entries tp[t] .. tp[t+1]-1 of tcol / tval, in stored order.
#define ZS_MIN_W 16
struct ZSell {
    int nsl = 0;
    int *so = nullptr, *rl = nullptr, *col = nullptr;
    zd *val = nullptr;
    int ntail = 0;
    int *trow = nullptr, *tp = nullptr, *tcol = nullptr;
    zd *tval = nullptr;
    // slice-diagonal copy (banded, ascending rows): slice offsets, per-slice column
    // offsets (position k of slice q at dso[q] / 64 + k), values (absent entries 0)
    long long dia = 0;
    int *dso = nullptr, *doff = nullptr;
    zd *dval = nullptr;
};

struct bkh_zmatrix {
    int n = 0;
    long long nnz = 0;
    ZSell A, AT;       // A and A^T (values stored unconjugated; A^H conjugates at use)
};

static void zsell_free(ZSell &s) {
    hipFree(s.so); hipFree(s.rl); hipFree(s.col); hipFree(s.val);
    hipFree(s.dso); hipFree(s.doff); hipFree(s.dval);
    hipFree(s.trow); hipFree(s.tp); hipFree(s.tcol); hipFree(s.tval);
    s = ZSell();
}

// The slice-diagonal copy of a complex matrix whose rows ascend, when it stays within
// 10 % of the nonzeros (BKH_DIA=0 turns it off): the solvers' SpMV reads per-slice
// column offsets instead of a column per entry. A value-0 position adds (0 + 0i) x,
// which leaves a finite row sum unchanged. Returns false only on a device error.
static bool zdia_upload(ZSell &s, int n, const std::vector<int> &rp, const std::vector<int> &ci,
                        const std::vector<zd> &v) {
    const char *env = getenv("BKH_DIA");
    if ((env && env[0] == '0') || rp[n] == 0) return true;
    for (int r = 0; r < n; r++)
        for (int k = rp[r] + 1; k < rp[r + 1]; k++)
            if (ci[k] <= ci[k - 1]) return true;
    const int nsl = s.nsl;
    std::vector<int> so((size_t)nsl + 1, 0), offs, u;
    long long tot = 0;
    const long long cap = (long long)(1.1 * (double)rp[n]) + 64;
    for (int q = 0; q < nsl; q++) {
        u.clear();
        for (int r = q * 64; r < std::min(n, q * 64 + 64); r++)
            for (int k = rp[r]; k < rp[r + 1]; k++) u.push_back(ci[k] - r);
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        tot += 64LL * (long long)u.size();
        if (tot > cap || tot >= (1LL << 31) - 64) return true;
        so[q + 1] = (int)tot;
        offs.insert(offs.end(), u.begin(), u.end());
    }
    std::vector<zd> dv((size_t)tot + 1, zmk(0.0, 0.0));
    for (int q = 0; q < nsl; q++) {
        const int *o = offs.data() + so[q] / 64;
        for (int r = q * 64; r < std::min(n, q * 64 + 64); r++) {
            int p = 0;
            for (int k = rp[r]; k < rp[r + 1]; k++) {
                while (o[p] != ci[k] - r) p++;
                dv[(size_t)so[q] + 64 * (size_t)p + (r & 63)] = v[k];
                p++;
            }
        }
    }
    offs.push_back(0);
    if (bkh_dmalloc(&s.dso, so.size() * sizeof(int)) != hipSuccess ||
        bkh_dmalloc(&s.doff, offs.size() * sizeof(int)) != hipSuccess ||
        bkh_dmalloc(&s.dval, dv.size() * sizeof(zd)) != hipSuccess) {
        bkh_set_error("hipMalloc failed for a complex matrix");
        return false;
    }
    if (!(hipMemcpy(s.dso, so.data(), so.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
          hipMemcpy(s.doff, offs.data(), offs.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
          hipMemcpy(s.dval, dv.data(), dv.size() * sizeof(zd), hipMemcpyHostToDevice) == hipSuccess))
        return false;
    s.dia = tot;
    return true;
}

static bool zsell_upload(ZSell &s, int n, const std::vector<int> &rp, const std::vector<int> &ci,
                         const std::vector<zd> &v) {
    const int nsl = (n + 63) / 64;
    std::vector<int> so((size_t)nsl + 1, 0), rl((size_t)std::max(n, 1), 0);
    std::vector<int> trow, tp(1, 0), tcol;
    std::vector<zd> tval;
    long long tot = 0;
    for (int q = 0; q < nsl; q++) {
        const int r0 = q * 64, r1 = std::min(n, q * 64 + 64);
        int w = 0;
        for (int r = r0; r < r1; r++) w = std::max(w, rp[r + 1] - rp[r]);
        const long long wcap = std::max<long long>(ZS_MIN_W, (2LL * (rp[r1] - rp[r0]) + 63) / 64);
        w = (int)std::min<long long>(w, wcap);
        for (int r = r0; r < r1; r++) {
            const int len = rp[r + 1] - rp[r];
            rl[r] = std::min(len, w);
            if (len > w) {
                trow.push_back(r);
                for (int k = rp[r] + w; k < rp[r + 1]; k++) { tcol.push_back(ci[k]); tval.push_back(v[k]); }
                tp.push_back((int)tcol.size());
            }
        }
        tot += 64LL * w;
        if (tot >= (1LL << 31) - 64) { bkh_set_error("complex matrix too large for 32-bit slice offsets"); return false; }
        so[q + 1] = (int)tot;
    }
    std::vector<int> col((size_t)std::max(tot, 1LL), 0);
    std::vector<zd> val((size_t)std::max(tot, 1LL), zmk(0.0, 0.0));
    for (int q = 0; q < nsl; q++)
        for (int r = q * 64; r < std::min(n, q * 64 + 64); r++) {
            const int w = (so[q + 1] - so[q]) / 64;
            for (int k = 0; k < w; k++) {
                const size_t e = (size_t)so[q] + 64 * (size_t)k + (r & 63);
                if (k < rl[r]) { col[e] = ci[rp[r] + k]; val[e] = v[rp[r] + k]; }
                else col[e] = r;
            }
        }
    s.nsl = nsl;
    if (bkh_dmalloc(&s.so, so.size() * sizeof(int)) != hipSuccess ||
        bkh_dmalloc(&s.rl, rl.size() * sizeof(int)) != hipSuccess ||
        bkh_dmalloc(&s.col, col.size() * sizeof(int)) != hipSuccess ||
        bkh_dmalloc(&s.val, val.size() * sizeof(zd)) != hipSuccess) {
        bkh_set_error("hipMalloc failed for a complex matrix");
        return false;
    }
    if (!(hipMemcpy(s.so, so.data(), so.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
          hipMemcpy(s.rl, rl.data(), rl.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
          hipMemcpy(s.col, col.data(), col.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
          hipMemcpy(s.val, val.data(), val.size() * sizeof(zd), hipMemcpyHostToDevice) == hipSuccess))
        return false;
    if (!trow.empty()) {
        if (bkh_dmalloc(&s.trow, trow.size() * sizeof(int)) != hipSuccess ||
            bkh_dmalloc(&s.tp, tp.size() * sizeof(int)) != hipSuccess ||
            bkh_dmalloc(&s.tcol, tcol.size() * sizeof(int)) != hipSuccess ||
            bkh_dmalloc(&s.tval, tval.size() * sizeof(zd)) != hipSuccess) {
            bkh_set_error("hipMalloc failed for a complex matrix");
            return false;
        }
        if (!(hipMemcpy(s.trow, trow.data(), trow.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(s.tp, tp.data(), tp.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(s.tcol, tcol.data(), tcol.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(s.tval, tval.data(), tval.size() * sizeof(zd), hipMemcpyHostToDevice) == hipSuccess))
            return false;
        s.ntail = (int)trow.size();
    }
    return zdia_upload(s, n, rp, ci, v);
}

extern "C" bkh_zmatrix *bkh_zmatrix_create(int n, const int *row_ptr, const int *col, const double *val) {
    if (n < 1 || !row_ptr || !col || !val) { bkh_set_error("bkh_zmatrix_create: bad arguments"); return nullptr; }
    if (!bkh_tctx()) return nullptr;
    const long long nnz = row_ptr[n];
    std::vector<int> rp(row_ptr, row_ptr + n + 1), ci(col, col + nnz);
    std::vector<zd> v((size_t)nnz);
    for (long long k = 0; k < nnz; k++) {
        if (ci[k] < 0 || ci[k] >= n) { bkh_set_error("bkh_zmatrix_create: column out of range"); return nullptr; }
        v[k] = zmk(val[2 * k], val[2 * k + 1]);
    }
    // A^T: a stable counting sort by column lists, for every output row c, the
    // scatter's contributions (src/matvec.c:85-90) in source-row order.
    std::vector<int> rpt((size_t)n + 1, 0), cit((size_t)nnz);
    std::vector<zd> vt((size_t)nnz);
    for (long long k = 0; k < nnz; k++) rpt[ci[k] + 1]++;
    for (int i = 0; i < n; i++) rpt[i + 1] += rpt[i];
    std::vector<int> fill(rpt.begin(), rpt.end() - 1);
    for (int j = 0; j < n; j++)
        for (int k = rp[j]; k < rp[j + 1]; k++) {
            const int d = fill[ci[k]]++;
            cit[d] = j;
            vt[d] = v[k];
        }
    bkh_zmatrix *A = new bkh_zmatrix();
    A->n = n;
    A->nnz = nnz;
    if (!zsell_upload(A->A, n, rp, ci, v) || !zsell_upload(A->AT, n, rpt, cit, vt)) {
        zsell_free(A->A); zsell_free(A->AT);
        delete A;
        return nullptr;
    }
    return A;
}

extern "C" long long bkh_zmatrix_dia_entries(const bkh_zmatrix *A) { return A ? A->A.dia + A->AT.dia : -1; }

extern "C" long long bkh_zmatrix_split_rows(const bkh_zmatrix *A) { return A ? A->A.ntail + A->AT.ntail : -1; }

extern "C" void bkh_zmatrix_destroy(bkh_zmatrix *A) {
    if (!A) return;
    hipDeviceSynchronize();
    zsell_free(A->A);
    zsell_free(A->AT);
    delete A;
    bkh_graph_epoch_bump();
}

// y = op(A) x, lane per row. CONJ: the A^T copy with conjugated values (A^H).
// DIA: on the slice-diagonal copy (so, col, val = its slice offsets, per-slice column
// offsets, values; the offsets are uniform over the wave's slice).
template <bool CONJ, bool DIA = false>
__global__ __launch_bounds__(256) void k_zspmv(int n, const int *__restrict__ so, const int *__restrict__ rl,
                                               const int *__restrict__ col, const zd *__restrict__ val,
                                               const zd *x, zd *y, const Ctl *c) {
    if (c && ctl_stopped(c)) return;
    const int row = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (row >= n) return;
    zd acc = zmk(0.0, 0.0);
    if constexpr (DIA) {
        const int sl = __builtin_amdgcn_readfirstlane(row >> 6);
        const int s0 = so[sl], w = (so[sl + 1] - s0) >> 6;
        const int *o = col + (s0 >> 6);
        const int base = s0 + (row & 63);
        for (int k = 0; k < w; k++) {
            zd a = val[base + 64 * k];
            if (CONJ) a = zconj(a);
            acc = zadd(acc, zmul(a, x[min(max(row + o[k], 0), n - 1)]));
        }
    } else {
        const int len = rl[row];
        const int base = so[row >> 6] + (row & 63);
        for (int k = 0; k < len; k++) {
            const int e = base + 64 * k;
            zd a = val[e];
            if (CONJ) a = zconj(a);
            acc = zadd(acc, zmul(a, x[col[e]]));
        }
    }
    y[row] = acc;
}

// The CSR tails of rows longer than their slice's SELL width, lane per such row: the
// sum the SELL kernel left in Y continues over the tail in stored order, so every row is
// summed entry by entry exactly as without the split. MM: the zcsrgemm1 form, each
// entry scaled by alpha before the product, for s columns.
template <bool CONJ, bool MM>
__global__ __launch_bounds__(256) void k_zsp_tail(int nt, int n, int s, const int *__restrict__ trow,
                                                  const int *__restrict__ tp, const int *__restrict__ tcol,
                                                  const zd *__restrict__ tval, zd alpha, const zd *X, zd *Y,
                                                  const Ctl *c) {
    if (c && ctl_stopped(c)) return;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nt) return;
    const int row = trow[t], k0 = tp[t], k1 = tp[t + 1];
    for (int j = 0; j < s; j++) {
        const zd *x = X + (size_t)j * n;
        zd acc = Y[(size_t)j * n + row];
        for (int k = k0; k < k1; k++) {
            zd a = tval[k];
            if (CONJ) a = zconj(a);
            if (MM) a = zmul(alpha, a);
            acc = zadd(acc, zmul(a, x[tcol[k]]));
        }
        Y[(size_t)j * n + row] = acc;
    }
}

template <bool MM>
static void launch_zsp_tail(hipStream_t st, const ZSell &M, bool herm, int n, int s, zd alpha, const zd *X, zd *Y,
                            const Ctl *c) {
    if (M.ntail == 0) return;
    const int G = (M.ntail + 255) / 256;
    if (herm)
        hipLaunchKernelGGL((k_zsp_tail<true, MM>), dim3(G), dim3(256), 0, st, M.ntail, n, s, M.trow, M.tp, M.tcol,
                           M.tval, alpha, X, Y, c);
    else
        hipLaunchKernelGGL((k_zsp_tail<false, MM>), dim3(G), dim3(256), 0, st, M.ntail, n, s, M.trow, M.tp,
                           M.tcol, M.tval, alpha, X, Y, c);
}

// dia: the solvers' own products (finite x) run on the slice-diagonal copy when the
// matrix has one; the public entry points keep SELL-64.
static void launch_zspmv(hipStream_t s, const bkh_zmatrix *A, bool herm, const zd *x, zd *y, const Ctl *c,
                         bool dia = true) {
    const ZSell &M = herm ? A->AT : A->A;
    const int G = (A->n + 255) / 256;
    if (dia && M.dia > 0) {
        if (herm)
            hipLaunchKernelGGL((k_zspmv<true, true>), dim3(G), dim3(256), 0, s, A->n, M.dso, M.rl, M.doff, M.dval,
                               x, y, c);
        else
            hipLaunchKernelGGL((k_zspmv<false, true>), dim3(G), dim3(256), 0, s, A->n, M.dso, M.rl, M.doff, M.dval,
                               x, y, c);
        return;
    }
    if (herm)
        hipLaunchKernelGGL((k_zspmv<true>), dim3(G), dim3(256), 0, s, A->n, M.so, M.rl, M.col, M.val, x, y, c);
    else
        hipLaunchKernelGGL((k_zspmv<false>), dim3(G), dim3(256), 0, s, A->n, M.so, M.rl, M.col, M.val, x, y, c);
    launch_zsp_tail<false>(s, M, herm, A->n, 1, zmk(1.0, 0.0), x, y, c);
}

// Y = beta Y + alpha op(A) X for s columns (column-major, ld n), the reference's
// zcsrgemm1 order: per row, beta*Y first, then + (alpha*a)*x per entry.
template <bool CONJ>
__global__ __launch_bounds__(256) void k_zspmm1(int n, int s, const int *__restrict__ so, const int *__restrict__ rl,
                                                const int *__restrict__ col, const zd *__restrict__ val,
                                                zd alpha, zd beta, const zd *X, zd *Y) {
    const int row = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (row >= n) return;
    const int len = rl[row];
    const int base = so[row >> 6] + (row & 63);
    for (int j = 0; j < s; j++) {
        const zd *x = X + (size_t)j * n;
        zd acc = zmul(beta, Y[(size_t)j * n + row]);
        for (int k = 0; k < len; k++) {
            const int e = base + 64 * k;
            zd a = val[e];
            if (CONJ) a = zconj(a);
            acc = zadd(acc, zmul(zmul(alpha, a), x[col[e]]));
        }
        Y[(size_t)j * n + row] = acc;
    }
}

// k_zspmm1 for s = S <= 8 known at compile time: each stored entry is loaded once and
// applied to all S columns (the runtime-s form re-reads A per column); per column every
// row still starts from beta Y and adds (alpha a) x in stored order, so Y is bitwise the
// same.
template <int S, bool CONJ>
__global__ __launch_bounds__(256) void k_zspmm1_t(int n, const int *__restrict__ so, const int *__restrict__ rl,
                                                  const int *__restrict__ col, const zd *__restrict__ val,
                                                  zd alpha, zd beta, const zd *X, zd *Y) {
    const int row = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (row >= n) return;
    const int len = rl[row];
    const int base = so[row >> 6] + (row & 63);
    zd acc[S];
#pragma unroll
    for (int j = 0; j < S; j++) acc[j] = zmul(beta, Y[(size_t)j * n + row]);
    for (int k = 0; k < len; k++) {
        const int e = base + 64 * k;
        zd a = val[e];
        if (CONJ) a = zconj(a);
        const zd aa = zmul(alpha, a);
        const int c = col[e];
#pragma unroll
        for (int j = 0; j < S; j++) acc[j] = zadd(acc[j], zmul(aa, X[(size_t)j * n + c]));
    }
#pragma unroll
    for (int j = 0; j < S; j++) Y[(size_t)j * n + row] = acc[j];
}

// Y = beta Y + alpha op(A) X over the SELL copy M (s columns); the long rows' CSR tail
// is added by launch_zsp_tail.
template <bool CONJ>
static void zspmm1_launch(hipStream_t st, const ZSell &M, int n, int s, zd al, zd be, const zd *X, zd *Y) {
    const int G = (n + 255) / 256;
    switch (s) {
#define ZSPMM_CASE(S_)                                                                                          \
    case S_:                                                                                                    \
        hipLaunchKernelGGL((k_zspmm1_t<S_, CONJ>), dim3(G), dim3(256), 0, st, n, M.so, M.rl, M.col, M.val, al, be, X, Y); \
        break;
        ZSPMM_CASE(1) ZSPMM_CASE(2) ZSPMM_CASE(3) ZSPMM_CASE(4)
        ZSPMM_CASE(5) ZSPMM_CASE(6) ZSPMM_CASE(7) ZSPMM_CASE(8)
#undef ZSPMM_CASE
    default:
        hipLaunchKernelGGL((k_zspmm1<CONJ>), dim3(G), dim3(256), 0, st, n, s, M.so, M.rl, M.col, M.val, al, be, X, Y);
    }
}

extern "C" int bkh_zspmv(const bkh_zmatrix *A, char op, const double *d_x, double *d_y) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !d_x || !d_y || (op != 'n' && op != 't')) return BKH_BAD_ARGUMENT;
    launch_zspmv(t->stream, A, op == 't', (const zd *)d_x, (zd *)d_y, nullptr, false);
    LAUNCH_OK();
    HIP_OK(hipStreamSynchronize(t->stream));
    return 0;
}

extern "C" int bkh_zspmm1(const bkh_zmatrix *A, char op, const double *alpha, int s, const double *d_X,
                          const double *beta, double *d_Y) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !alpha || !beta || !d_X || !d_Y || s < 1 || (op != 'n' && op != 't')) return BKH_BAD_ARGUMENT;
    const ZSell &M = op == 't' ? A->AT : A->A;
    const int G = (A->n + 255) / 256;
    const zd al = zmk(alpha[0], alpha[1]), be = zmk(beta[0], beta[1]);
    (void)G;
    if (op == 't') zspmm1_launch<true>(t->stream, M, A->n, s, al, be, (const zd *)d_X, (zd *)d_Y);
    else zspmm1_launch<false>(t->stream, M, A->n, s, al, be, (const zd *)d_X, (zd *)d_Y);
    launch_zsp_tail<true>(t->stream, M, op == 't', A->n, s, al, (const zd *)d_X, (zd *)d_Y, nullptr);
    LAUNCH_OK();
    HIP_OK(hipStreamSynchronize(t->stream));
    return 0;
}

// ------------------------------------------------------------- dot products

// Up to ZD_MAX conj(x_d).y_d sums per launch; partial (re, im) pairs per workgroup.
#define ZD_MAX 4
struct ZDots {
    const zd *x[ZD_MAX];
    const zd *y[ZD_MAX];
};

template <int ND>
__global__ __launch_bounds__(256) void k_zdots(long long n, ZDots a, const Ctl *c, double *part) {
    if (c && ctl_stopped(c)) return;
    double v[2 * ND];
#pragma unroll
    for (int d = 0; d < 2 * ND; d++) v[d] = 0.0;
    ZLOOP(i, n) {
#pragma unroll
        for (int d = 0; d < ND; d++) {
            const zd p = zmul(zconj(a.x[d][i]), a.y[d][i]);
            v[2 * d] += p.x;
            v[2 * d + 1] += p.y;
        }
    }
    block_partial<2 * ND, 256>(v, part);
}

template <int ND>
static void launch_zdots(hipStream_t s, long long n, const ZDots &a, const Ctl *c, double *part) {
    hipLaunchKernelGGL((k_zdots<ND>), dim3(vec_grid(n, 256)), dim3(256), 0, s, n, a, c, part);
}

static ZDots zdots(const zd *x0, const zd *y0, const zd *x1 = nullptr, const zd *y1 = nullptr) {
    ZDots a;
    memset(&a, 0, sizeof a);
    a.x[0] = x0; a.y[0] = y0; a.x[1] = x1; a.y[1] = y1;
    return a;
}

// ------------------------------------------------------------------ solvers

// Ctl::s slots (complex values take two doubles).
enum {
    ZS_ALPHA = 0, ZS_BETA = 2, ZS_RDOT = 4, ZS_G0 = 6, ZS_G1 = 8, ZS_ALOLD = 10, ZS_RHOOLD = 12,
    ZS_OMEGA = 20,      // omega[0..l]        (up to 17 values)
    ZS_OMD = 56,        // omega_d[0..l]
    ZS_OMDD = 92,       // omega_dd[0..l]
    ZS_SIG = 128,       // sigma[0..l] (MR norms, real parts used)
};
#define ZS_LMAX 8

static inline int zbl_ok(int l) { return l >= 1 && l <= ZS_LMAX; }

// r = -(A x) + 1 b (zdscal(-1) then zaxpy(one), bicg.c:106-109); copies to the
// shadow / direction vectors; partial conj(sr).r and |b|^2.
__global__ __launch_bounds__(256) void k_zinit(long long n, const zd *b, const zd *ax, zd *r, zd *c1, zd *c2, zd *c3,
                                               double *part) {
    double v[3] = {0.0, 0.0, 0.0};
    ZLOOP(i, n) {
        const zd t = zmk(-1.0 * ax[i].x, -1.0 * ax[i].y);
        const zd ri = zadd(t, zmul(zmk(1.0, 0.0), b[i]));
        r[i] = ri;
        if (c1) c1[i] = ri;
        if (c2) c2[i] = ri;
        if (c3) c3[i] = ri;
        const zd d = zmul(zconj(ri), ri);
        v[0] += d.x;
        v[1] += d.y;
        v[2] += b[i].x * b[i].x + b[i].y * b[i].y;
    }
    block_partial<3, 256>(v, part);
}

// |r0|^2 of the initial residual and |b|^2 -> Ctl::fnb; slot RDOT = conj(r).r.
struct ZInitFin {
    __device__ void operator()(Ctl *c, const double *tot) const {
        c->s[ZS_RDOT] = tot[0];
        c->s[ZS_RDOT + 1] = tot[1];
        c->fnb = sqrt(tot[2]);
        if (!(c->fnb > 0.0)) ctl_breakdown(c);
    }
};

// ---- BiCG (bicg.c:91-177)

struct ZbgAlphaFin {   // alpha = r_dot_old / conj(sp).ap   (:125-126)
    __device__ void operator()(Ctl *c, const double *tot) const {
        const zd d = zmk(tot[0], tot[1]);
        if (zbad(d)) { ctl_breakdown(c); return; }
        zstore(c->s, ZS_ALPHA / 2, zdiv(zload(c->s, ZS_RDOT / 2), d));
    }
};

// x += alpha p; r += (-alpha) ap; sr += (-conj(alpha)) atp; partials conj(sr).r, |r|^2.
__global__ __launch_bounds__(256) void k_zbg_upd(long long n, zd *x, zd *r, zd *sr, const zd *p, const zd *ap,
                                                 const zd *atp, const Ctl *c, double *part) {
    if (ctl_stopped(c)) return;
    const zd al = zload(c->s, ZS_ALPHA / 2), mal = zneg(al), mcal = zneg(zconj(al));
    double v[3] = {0.0, 0.0, 0.0};
    ZLOOP(i, n) {
        x[i] = zadd(x[i], zmul(al, p[i]));
        const zd ri = zadd(r[i], zmul(mal, ap[i]));
        r[i] = ri;
        const zd si = zadd(sr[i], zmul(mcal, atp[i]));
        sr[i] = si;
        const zd d = zmul(zconj(si), ri);
        v[0] += d.x;
        v[1] += d.y;
        v[2] += ri.x * ri.x + ri.y * ri.y;
    }
    block_partial<3, 256>(v, part);
}

struct ZbgBetaFin {    // beta = r_dot / r_dot_old (:141-144); error test (:156-159)
    __device__ void operator()(Ctl *c, const double *tot) const {
        const zd rd = zmk(tot[0], tot[1]), old = zload(c->s, ZS_RDOT / 2);
        zstore(c->s, ZS_BETA / 2, zdiv(rd, old));
        zstore(c->s, ZS_RDOT / 2, rd);
        ctl_check(c, tot[2]);
    }
};

// p = beta p + 1 r; sp = conj(beta) sp + 1 sr (zscal then zaxpy, :146-153).
__global__ __launch_bounds__(256) void k_zbg_p(long long n, zd *p, zd *sp, const zd *r, const zd *sr, const Ctl *c) {
    if (ctl_stopped(c)) return;
    const zd be = zload(c->s, ZS_BETA / 2), cbe = zconj(be), one = zmk(1.0, 0.0);
    ZLOOP(i, n) {
        p[i] = zadd(zmul(be, p[i]), zmul(one, r[i]));
        sp[i] = zadd(zmul(cbe, sp[i]), zmul(one, sr[i]));
    }
}

struct ZResFin {
    __device__ void operator()(Ctl *c, const double *tot) const {
        c->s[0] = tot[1] > 0.0 ? sqrt(tot[0]) / sqrt(tot[1]) : 0.0;
    }
};

// ||b - (A + sigma I) x|| / ||b|| on the device into co->s[0] (co: a scratch control block).
__global__ __launch_bounds__(256) void k_zres2_shift(long long n, const zd *b, const zd *ax, const zd *x, zd sigma,
                                                     double *part) {
    double v[2] = {0.0, 0.0};
    ZLOOP(i, n) {
        const zd r = zsub(zsub(b[i], ax[i]), zmul(sigma, x[i]));
        v[0] += r.x * r.x + r.y * r.y;
        v[1] += b[i].x * b[i].x + b[i].y * b[i].y;
    }
    block_partial<2, 256>(v, part);
}

static int ztrue_res(ThreadCtx *t, const bkh_zmatrix *A, const zd *b, const zd *x, zd sigma, zd *tmp, Ctl *co,
                     double *part, double *out) {
    hipStream_t s = t->stream;
    const long long n = A->n;
    if (ctl_init(t, co, 0.0, 1)) return BKH_DEVICE_ERROR;
    launch_zspmv(s, A, false, x, tmp, nullptr);
    hipLaunchKernelGGL(k_zres2_shift, dim3(vec_grid(n, 256)), dim3(256), 0, s, n, b, (const zd *)tmp, x, sigma, part);
    launch_finalize<2>(s, part, vec_grid(n, 256), co, ZResFin());
    LAUNCH_OK();
    if (ctl_fetch(t, co)) return BKH_DEVICE_ERROR;
    *out = t->ctl_host->s[0];
    return 0;
}

// Fills `inf` from the solve's control block (fetched into t->ctl_host by the caller).
static int zresult(ThreadCtx *t, const Ctl *hc, bkh_info *inf) {
    inf->status = hc->status;
    inf->iters = hc->result_it;
    inf->rec_res = hc->err;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t->ev0, t->ev1) == hipSuccess) inf->time_ms = ms;
    bkh_note_solve_ms(ms);
    inf->ret = (hc->status == ST_CONVERGED) ? hc->result_it : (hc->status == ST_BREAKDOWN ? -2 : -1);
    return inf->ret;
}

static int zsolve_end(ThreadCtx *t, const bkh_zmatrix *A, const zd *b, const zd *x, Ctl *c, Ctl *co, zd *tmp,
                      double *part, bkh_info *info) {
    HIP_OK(hipEventRecord(t->ev1, t->stream));
    if (ctl_fetch(t, c)) return BKH_DEVICE_ERROR;
    bkh_info inf;
    memset(&inf, 0, sizeof inf);
    const Ctl hc = *t->ctl_host;
    HIP_OK(hipEventSynchronize(t->ev1));
    int ret = zresult(t, &hc, &inf);
    if (info) {
        if (ztrue_res(t, A, b, x, zmk(0.0, 0.0), tmp, co, part, &inf.true_res)) return BKH_DEVICE_ERROR;
        *info = inf;
    }
    return ret;
}

extern "C" int bkh_zbicg(const bkh_zmatrix *A, const double *d_b, double *d_x, double tol, int max_iter,
                         bkh_info *info) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !d_b || !d_x) return BKH_BAD_ARGUMENT;
    if (max_iter <= 0) {                 // no iteration: the reference's error test sees 0
        if (info) { memset(info, 0, sizeof *info); info->status = ST_CONVERGED; }
        return 0;
    }
    const long long n = A->n;
    const int G = vec_grid(n, 256);
    void *ws = bkh_workspace(2 * carve_size(sizeof(Ctl)) + carve_size(sizeof(double) * 6 * G) +
                             6 * carve_size(sizeof(zd) * n));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1), *co = cv.take<Ctl>(1);
    double *part = cv.take<double>(6 * G);
    zd *r = cv.take<zd>(n), *sr = cv.take<zd>(n), *p = cv.take<zd>(n), *sp = cv.take<zd>(n);
    zd *ap = cv.take<zd>(n), *atp = cv.take<zd>(n);
    const zd *b = (const zd *)d_b;
    zd *x = (zd *)d_x;
    hipStream_t s = t->stream;
    HIP_OK(hipEventRecord(t->ev0, s));
    if (ctl_init(t, c, tol, max_iter)) return BKH_DEVICE_ERROR;
    launch_zspmv(s, A, false, x, ap, nullptr);                 // bicg.c:106-115
    hipLaunchKernelGGL(k_zinit, dim3(G), dim3(256), 0, s, n, b, (const zd *)ap, r, sr, p, sp, part);
    launch_finalize<3>(s, part, G, c, ZInitFin());
    LAUNCH_OK();
    const int K = default_graph_iters(n);
    hipGraphExec_t ge = capture_cached(t, {0x2B16, (uintptr_t)A, (uintptr_t)b, (uintptr_t)x, (uintptr_t)c,
                                           (uintptr_t)part, (uintptr_t)r}, s, K, [=]() {
        launch_zspmv(s, A, false, p, ap, c);
        launch_zspmv(s, A, true, sp, atp, c);
        launch_zdots<1>(s, n, zdots(sp, ap), c, part);
        launch_finalize<2>(s, part, G, c, ZbgAlphaFin());
        hipLaunchKernelGGL(k_zbg_upd, dim3(G), dim3(256), 0, s, n, x, r, sr, (const zd *)p, (const zd *)ap,
                           (const zd *)atp, (const Ctl *)c, part);
        launch_finalize<3>(s, part, G, c, ZbgBetaFin());
        hipLaunchKernelGGL(k_zbg_p, dim3(G), dim3(256), 0, s, n, p, sp, (const zd *)r, (const zd *)sr,
                           (const Ctl *)c);
    });
    if (!ge) { bkh_set_error("hipGraph capture of the complex BiCG iteration failed"); return BKH_DEVICE_ERROR; }
    if (run_graph_until_done(t, ge, c, max_iter / K + 2)) return BKH_DEVICE_ERROR;
    return zsolve_end(t, A, b, x, c, co, ap, part, info);
}

// ---- BiCGSTAB(l) (bicgstabl.c:141-276)
//
// Vectors: R[0..l], P[0..l] (n each), rs; scalars in Ctl::s (complex pairs).

// BiCG step j, first half: gamma1 = conj(rs).r_j; beta = (-alpha) gamma1 / gamma0;
// gamma0 = gamma1 (:189-191). On j = 0 the outer gamma0 = -omega_l gamma0 (:186)
// comes first.
struct ZblBetaFin {
    int j, l;
    __device__ void operator()(Ctl *c, const double *tot) const {
        if (j == 0) zstore(c->s, ZS_G0 / 2, zmul(zneg(zload(c->s, ZS_OMEGA / 2 + l)), zload(c->s, ZS_G0 / 2)));
        const zd g1 = zmk(tot[0], tot[1]), g0 = zload(c->s, ZS_G0 / 2);
        if (zbad(g0)) { ctl_breakdown(c); return; }
        zstore(c->s, ZS_BETA / 2, zdiv(zmul(zneg(zload(c->s, ZS_ALPHA / 2)), g1), g0));
        zstore(c->s, ZS_G0 / 2, g1);
        zstore(c->s, ZS_G1 / 2, g1);
    }
};

// p_i = beta p_i + 1 r_i, i <= j (:193-196).
__global__ __launch_bounds__(256) void k_zbl_p(long long n, int j, zd *P, const zd *R, const Ctl *c) {
    if (ctl_stopped(c)) return;
    const zd be = zload(c->s, ZS_BETA / 2), one = zmk(1.0, 0.0);
    ZLOOP(e, n) {
        for (int i = 0; i <= j; i++) {
            const size_t k = (size_t)i * n + e;
            P[k] = zadd(zmul(be, P[k]), zmul(one, R[k]));
        }
    }
}

struct ZblAlphaFin {   // alpha = gamma1 / conj(rs).p_{j+1} (:200-201)
    __device__ void operator()(Ctl *c, const double *tot) const {
        const zd d = zmk(tot[0], tot[1]);
        if (zbad(d)) { ctl_breakdown(c); return; }
        zstore(c->s, ZS_ALPHA / 2, zdiv(zload(c->s, ZS_G1 / 2), d));
    }
};

// r_i += (-alpha) p_{i+1}, i <= j (:203-206); x += alpha p_0 (:209).
__global__ __launch_bounds__(256) void k_zbl_r(long long n, int j, zd *R, const zd *P, zd *x, const Ctl *c) {
    if (ctl_stopped(c)) return;
    const zd al = zload(c->s, ZS_ALPHA / 2), mal = zneg(al);
    ZLOOP(e, n) {
        for (int i = 0; i <= j; i++) {
            const size_t k = (size_t)i * n + e;
            R[k] = zadd(R[k], zmul(mal, P[k + n]));
        }
        x[e] = zadd(x[e], zmul(al, P[e]));
    }
}

// MR part, modified Gram-Schmidt (:214-224). tau and sigma indices follow the
// reference (tau stride l). Slot layout: tau in `tau` (device array, l2*l2 pairs).
struct ZblTauFin {     // tau[j*l+i] = conj(r_i).r_j / sigma_i
    int i, j, l;
    double *tau;
    __device__ void operator()(Ctl *c, const double *tot) const {
        const zd sg = zload(c->s, ZS_SIG / 2 + i);
        if (zbad(sg)) { ctl_breakdown(c); return; }
        zstore(tau, j * l + i, zdiv(zmk(tot[0], tot[1]), sg));
    }
};

// r_j += (-tau[j*l+i]) r_i.
__global__ __launch_bounds__(256) void k_zbl_mgs(long long n, zd *rj, const zd *ri, const double *tau, int q,
                                                 const Ctl *c) {
    if (ctl_stopped(c)) return;
    const zd m = zneg(zload(tau, q));
    ZLOOP(e, n) rj[e] = zadd(rj[e], zmul(m, ri[e]));
}

// sigma_j = conj(r_j).r_j; omega_d[j] = conj(r_j).r_0 / sigma_j; on j == l the
// omega / omega_dd recurrences (:226-242).
struct ZblSigFin {
    int j, l;
    const double *tau;
    __device__ void operator()(Ctl *c, const double *tot) const {
        const zd sg = zmk(tot[0], tot[1]);
        if (zbad(sg)) { ctl_breakdown(c); return; }
        zstore(c->s, ZS_SIG / 2 + j, sg);
        zstore(c->s, ZS_OMD / 2 + j, zdiv(zmk(tot[2], tot[3]), sg));
        if (j < l) return;
        zd *om = (zd *)&c->s[ZS_OMEGA], *omd = (zd *)&c->s[ZS_OMD], *omdd = (zd *)&c->s[ZS_OMDD];
        om[l] = omd[l];
        for (int jj = l - 1; jj > 0; jj--) {
            zd t = zmk(0.0, 0.0);
            for (int i = jj + 1; i <= l; i++) t = zadd(t, zmul(zload(tau, i * l + jj), om[i]));
            om[jj] = zsub(omd[jj], t);
        }
        for (int jj = 1; jj < l; jj++) {
            zd t = zmk(0.0, 0.0);
            for (int i = jj + 1; i < l; i++) t = zadd(t, zmul(zload(tau, i * l + jj), om[i + 1]));
            omdd[jj] = zadd(om[jj + 1], t);
        }
    }
};

// Update (:245-256): x += omega_1 r_0, r_0 += (-omega_d_l) r_l, p_0 += (-omega_l) p_l,
// then for i = 1..l-1: p_0 += (-omega_i) p_i, r_0 += (-omega_d_i) r_i, x += omega_dd_i r_i;
// partial |r_0|^2 for the error test.
__global__ __launch_bounds__(256) void k_zbl_upd(long long n, int l, zd *R, zd *P, zd *x, const Ctl *c, double *part) {
    if (ctl_stopped(c)) return;
    const zd *om = (const zd *)&c->s[ZS_OMEGA], *omd = (const zd *)&c->s[ZS_OMD], *omdd = (const zd *)&c->s[ZS_OMDD];
    double v[1] = {0.0};
    ZLOOP(e, n) {
        zd xe = zadd(x[e], zmul(om[1], R[e]));
        zd r0 = zadd(R[e], zmul(zneg(omd[l]), R[(size_t)l * n + e]));
        zd p0 = zadd(P[e], zmul(zneg(om[l]), P[(size_t)l * n + e]));
        for (int i = 1; i < l; i++) {
            const zd ri = R[(size_t)i * n + e];
            p0 = zadd(p0, zmul(zneg(om[i]), P[(size_t)i * n + e]));
            r0 = zadd(r0, zmul(zneg(omd[i]), ri));
            xe = zadd(xe, zmul(omdd[i], ri));
        }
        x[e] = xe;
        R[e] = r0;
        P[e] = p0;
        v[0] += r0.x * r0.x + r0.y * r0.y;
    }
    block_partial<1, 256>(v, part);
}

struct ZCheckFin {
    __device__ void operator()(Ctl *c, const double *tot) const { ctl_check(c, tot[0]); }
};

__global__ void k_zbl_setup(Ctl *c, int l) {
    zstore(c->s, ZS_OMEGA / 2 + l, zmk(1.0, 0.0));      // omega[l] = 1 (:168)
    zstore(c->s, ZS_G0 / 2, zmk(1.0, 0.0));             // gamma0 = 1
    zstore(c->s, ZS_ALPHA / 2, zmk(0.0, 0.0));          // alpha = 0
}

extern "C" int bkh_zbicgstabl(const bkh_zmatrix *A, const double *d_b, int l, double *d_x, double tol,
                              int max_iter, bkh_info *info) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !d_b || !d_x || !zbl_ok(l)) return BKH_BAD_ARGUMENT;
    if (max_iter <= 0) {                 // no iteration: the reference's error test sees 0
        if (info) { memset(info, 0, sizeof *info); info->status = ST_CONVERGED; }
        return 0;
    }
    const long long n = A->n;
    const int G = vec_grid(n, 256), l2 = l + 1;
    void *ws = bkh_workspace(2 * carve_size(sizeof(Ctl)) + carve_size(sizeof(double) * 8 * G) +
                             carve_size(sizeof(zd) * l2 * l2) + (2 * l2 + 2) * carve_size(sizeof(zd) * n));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1), *co = cv.take<Ctl>(1);
    double *part = cv.take<double>(8 * G);
    double *tau = (double *)cv.take<zd>((size_t)l2 * l2);
    zd *R = cv.take<zd>((size_t)l2 * n), *P = cv.take<zd>((size_t)l2 * n), *rs = cv.take<zd>(n), *tmp = cv.take<zd>(n);
    const zd *b = (const zd *)d_b;
    zd *x = (zd *)d_x;
    hipStream_t s = t->stream;
    HIP_OK(hipEventRecord(t->ev0, s));
    if (ctl_init(t, c, tol, max_iter)) return BKH_DEVICE_ERROR;
    HIP_OK(hipMemsetAsync(P, 0, sizeof(zd) * l2 * n, s));
    HIP_OK(hipMemsetAsync(R, 0, sizeof(zd) * l2 * n, s));
    hipLaunchKernelGGL(k_zbl_setup, dim3(1), dim3(1), 0, s, c, l);
    launch_zspmv(s, A, false, x, tmp, nullptr);                // :175-181
    hipLaunchKernelGGL(k_zinit, dim3(G), dim3(256), 0, s, n, b, (const zd *)tmp, R, P, rs, (zd *)nullptr, part);
    launch_finalize<3>(s, part, G, c, ZInitFin());
    LAUNCH_OK();
    hipGraphExec_t ge = capture_cached(t, {0x2B17, (uintptr_t)A, (uintptr_t)b, (uintptr_t)x, (uintptr_t)c,
                                           (uintptr_t)part, (uintptr_t)R, (uintptr_t)l}, s, 1, [=]() {
        for (int j = 0; j < l; j++) {
            launch_zdots<1>(s, n, zdots(rs, R + (size_t)j * n), c, part);
            launch_finalize<2>(s, part, G, c, ZblBetaFin{j, l});
            hipLaunchKernelGGL(k_zbl_p, dim3(G), dim3(256), 0, s, n, j, P, (const zd *)R, (const Ctl *)c);
            launch_zspmv(s, A, false, P + (size_t)j * n, P + (size_t)(j + 1) * n, c);
            launch_zdots<1>(s, n, zdots(rs, P + (size_t)(j + 1) * n), c, part);
            launch_finalize<2>(s, part, G, c, ZblAlphaFin());
            hipLaunchKernelGGL(k_zbl_r, dim3(G), dim3(256), 0, s, n, j, R, (const zd *)P, x, (const Ctl *)c);
            launch_zspmv(s, A, false, R + (size_t)j * n, R + (size_t)(j + 1) * n, c);
        }
        for (int j = 1; j <= l; j++) {
            zd *rj = R + (size_t)j * n;
            for (int i = 1; i < j; i++) {
                const zd *ri = R + (size_t)i * n;
                launch_zdots<1>(s, n, zdots(ri, rj), c, part);
                launch_finalize<2>(s, part, G, c, ZblTauFin{i, j, l, tau});
                hipLaunchKernelGGL(k_zbl_mgs, dim3(G), dim3(256), 0, s, n, rj, ri, (const double *)tau, j * l + i,
                                   (const Ctl *)c);
            }
            launch_zdots<2>(s, n, zdots(rj, rj, rj, R), c, part);
            launch_finalize<4>(s, part, G, c, ZblSigFin{j, l, tau});
        }
        hipLaunchKernelGGL(k_zbl_upd, dim3(G), dim3(256), 0, s, n, l, R, P, x, (const Ctl *)c, part);
        launch_finalize<1>(s, part, G, c, ZCheckFin());
    });
    if (!ge) { bkh_set_error("hipGraph capture of the complex BiCGSTAB(l) iteration failed"); return BKH_DEVICE_ERROR; }
    if (run_graph_until_done(t, ge, c, max_iter + 2)) return BKH_DEVICE_ERROR;
    return zsolve_end(t, A, b, x, c, co, tmp, part, info);
}

// ---- shifted BiCGSTAB(l) (shbicgstabl.c:288-592)
//
// Per-shift scalars in a device array (one thread per shift in the scalar kernels);
// p_shift[k][i] at PS + (k (l+1) + i) n, x_k at x + (k+1) n as in the reference;
// the per-shift vector kernels run a 2D grid (rows x shifts).
struct ZSh {
    zd sigma, pi, pi_new, pi_old, theta, alpha_s, beta_s;
    zd cA, cB, cC, d1, d2;           // coefficients of the BiCG step's p_shift updates
    zd cx[ZS_LMAX];                  // x_k coefficients of the MR update
    zd cpsi;                         // 1 / psi
    zd ck[ZS_LMAX][6];               // every step's cA, cB, cC, d1, d2, alpha_s (fused pass)
};

__global__ void k_zsh_setup(Ctl *c, ZSh *sh, const double *sigma, int ns, int l) {
    if (threadIdx.x == 0) {
        zstore(c->s, ZS_OMEGA / 2 + l - 1, zmk(1.0, 0.0));   // omega[l-1] = 1 (:315)
        zstore(c->s, ZS_ALPHA / 2, zmk(0.0, 0.0));           // alpha = 0, alpha_old = 1,
        zstore(c->s, ZS_ALOLD / 2, zmk(1.0, 0.0));           // rho_old = 1 (:330-332)
        zstore(c->s, ZS_RHOOLD / 2, zmk(1.0, 0.0));
    }
    for (int k = threadIdx.x; k < ns; k += blockDim.x) {
        ZSh z;
        memset(&z, 0, sizeof z);
        z.sigma = zmk(sigma[2 * k], sigma[2 * k + 1]);
        z.pi = z.pi_old = z.theta = zmk(1.0, 0.0);          // :366-370
        sh[k] = z;
    }
}

// Seed step j: rho = conj(rs).r_j, beta = alpha rho / rho_old (:390-391); on j = 0
// first rho_old = -omega[l-1] rho_old (:387).
struct ZshBetaFin {
    int j, l;
    __device__ void operator()(Ctl *c, const double *tot) const {
        if (j == 0)
            zstore(c->s, ZS_RHOOLD / 2, zmul(zneg(zload(c->s, ZS_OMEGA / 2 + l - 1)), zload(c->s, ZS_RHOOLD / 2)));
        const zd rho = zmk(tot[0], tot[1]), ro = zload(c->s, ZS_RHOOLD / 2);
        if (zbad(ro)) { ctl_breakdown(c); return; }
        zstore(c->s, ZS_BETA / 2, zdiv(zmul(zload(c->s, ZS_ALPHA / 2), rho), ro));
        zstore(c->s, ZS_G1 / 2, rho);
    }
};

// p_i = (-beta) p_i + 1 r_i, i <= j (:394-398).
__global__ __launch_bounds__(256) void k_zsh_p(long long n, int j, zd *P, const zd *R, const Ctl *c) {
    if (ctl_stopped(c)) return;
    const zd mb = zneg(zload(c->s, ZS_BETA / 2)), one = zmk(1.0, 0.0);
    ZLOOP(e, n) {
        for (int i = 0; i <= j; i++) {
            const size_t k = (size_t)i * n + e;
            P[k] = zadd(zmul(mb, P[k]), zmul(one, R[k]));
        }
    }
}

struct ZshAlphaFin {   // alpha = rho / conj(rs).p_{j+1} (:405-406)
    __device__ void operator()(Ctl *c, const double *tot) const {
        const zd d = zmk(tot[0], tot[1]);
        if (zbad(d)) { ctl_breakdown(c); return; }
        zstore(c->s, ZS_ALPHA / 2, zdiv(zload(c->s, ZS_G1 / 2), d));
    }
};

// x += alpha p_0 (:409).
__global__ __launch_bounds__(256) void k_zsh_x(long long n, zd *x, const zd *P, const Ctl *c) {
    if (ctl_stopped(c)) return;
    const zd al = zload(c->s, ZS_ALPHA / 2);
    ZLOOP(e, n) x[e] = zadd(x[e], zmul(al, P[e]));
}

// Per-shift scalars of step j (:414-424) and the coefficients of both p_shift updates.
__global__ void k_zsh_sc1(int ns, ZSh *sh, const Ctl *c, int j) {
    if (ctl_stopped(c)) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    ZSh &z = sh[k];   // in place: a copy of the whole state spills to scratch
    const zd al = zload(c->s, ZS_ALPHA / 2), be = zload(c->s, ZS_BETA / 2), ao = zload(c->s, ZS_ALOLD / 2);
    const zd one = zmk(1.0, 0.0);
    const zd as = zmul(al, z.sigma);
    const zd t1 = zmul(zmk(1.0 + as.x, as.y), z.pi);
    const zd t2 = zmul(zdiv(zmul(al, be), ao), zsub(z.pi_old, z.pi));
    z.pi_new = zadd(t1, t2);
    const zd q = zdiv(z.pi_old, z.pi);
    z.beta_s = zmul(zmul(q, q), be);
    z.alpha_s = zmul(zdiv(z.pi, z.pi_new), al);
    z.cA = zneg(z.beta_s);
    z.cB = zdiv(one, zmul(z.pi, z.theta));
    z.cC = zdiv(one, z.pi);
    z.d1 = zdiv(zmk(-1.0, 0.0), z.pi_new);
    z.d2 = zdiv(one, zmul(z.theta, z.alpha_s));
    z.ck[j][0] = z.cA; z.ck[j][1] = z.cB; z.ck[j][2] = z.cC;
    z.ck[j][3] = z.d1; z.ck[j][4] = z.d2; z.ck[j][5] = z.alpha_s;
}

// p_shift updates of step j before the seed's r update (:426-437): p_s,i = cA p_s,i +
// cB r_i (i <= j); p_s,j+1 = cC r_j; x_k += alpha_s p_s,0.
__global__ __launch_bounds__(256) void k_zsh_v1(long long n, int j, int l2, zd *PS, zd *XS, const zd *R,
                                                const ZSh *sh, const Ctl *c) {
    if (ctl_stopped(c)) return;
    const int k = blockIdx.y;
    const ZSh z = sh[k];
    zd *ps = PS + (size_t)k * l2 * n, *xk = XS + (size_t)(k + 1) * n;
    ZLOOP(e, n) {
        for (int i = 0; i <= j; i++) {
            const size_t q = (size_t)i * n + e;
            ps[q] = zadd(zmul(z.cA, ps[q]), zmul(z.cB, R[q]));
        }
        ps[(size_t)(j + 1) * n + e] = zmul(z.cC, R[(size_t)j * n + e]);
        xk[e] = zadd(xk[e], zmul(z.alpha_s, ps[e]));
    }
}

// r_i += (-alpha) p_{i+1}, i <= j (:443-446).
// snap (fused per-shift pass): r_i before this update to snap[j(j+1)/2 + i], and on the
// last step (j = l - 1) r_j after it to snap[l(l+1)/2].
__global__ __launch_bounds__(256) void k_zsh_r(long long n, int j, zd *R, const zd *P, const Ctl *c, zd *snap,
                                               int l) {
    if (ctl_stopped(c)) return;
    const zd mal = zneg(zload(c->s, ZS_ALPHA / 2));
    zd *sj = snap ? snap + (size_t)(j * (j + 1) / 2) * n : nullptr;
    ZLOOP(e, n) {
        for (int i = 0; i <= j; i++) {
            const size_t k = (size_t)i * n + e;
            const zd r = R[k];
            if (sj) sj[(size_t)i * n + e] = r;
            R[k] = zadd(r, zmul(mal, P[k + n]));
        }
        if (sj && j == l - 1) snap[(size_t)(l * (l + 1) / 2) * n + e] = R[(size_t)j * n + e];
    }
}

// p_s,j+1 = d2 (p_s,j+1 + d1 r_j) + (-sigma) p_s,j (:452-461).
__global__ __launch_bounds__(256) void k_zsh_v2(long long n, int j, int l2, zd *PS, const zd *R, const ZSh *sh,
                                                const Ctl *c) {
    if (ctl_stopped(c)) return;
    const int k = blockIdx.y;
    const ZSh z = sh[k];
    const zd ms = zneg(z.sigma);
    zd *ps = PS + (size_t)k * l2 * n;
    ZLOOP(e, n) {
        const size_t q = (size_t)(j + 1) * n + e;
        zd v = zadd(ps[q], zmul(z.d1, R[(size_t)j * n + e]));
        v = zmul(z.d2, v);
        ps[q] = zadd(v, zmul(ms, ps[(size_t)j * n + e]));
    }
}

// alpha_old = alpha, rho_old = rho; pi_old = pi, pi = pi_new (:464-469).
__global__ void k_zsh_rot(int ns, ZSh *sh, Ctl *c) {
    if (ctl_stopped(c)) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {
        zstore(c->s, ZS_ALOLD / 2, zload(c->s, ZS_ALPHA / 2));
        zstore(c->s, ZS_RHOOLD / 2, zload(c->s, ZS_G1 / 2));
    }
    if (k < ns) {
        sh[k].pi_old = sh[k].pi;
        sh[k].pi = sh[k].pi_new;
    }
}

// MR part (:475-502): tau[j*l+i] = conj(r_{i+1}).r_{j+1} / s_i; s_j; omega_d_j; and on
// the last j the omega / omega_dd recurrences.
struct ZshSigFin {
    int j, l;
    const double *tau;
    __device__ void operator()(Ctl *c, const double *tot) const {
        const zd sg = zmk(tot[0], tot[1]);
        if (zbad(sg)) { ctl_breakdown(c); return; }
        zstore(c->s, ZS_SIG / 2 + j, sg);
        zstore(c->s, ZS_OMD / 2 + j, zdiv(zmk(tot[2], tot[3]), sg));
        if (j < l - 1) return;
        zd *om = (zd *)&c->s[ZS_OMEGA], *omd = (zd *)&c->s[ZS_OMD], *omdd = (zd *)&c->s[ZS_OMDD];
        om[l - 1] = omd[l - 1];
        for (int jj = l - 2; jj >= 0; jj--) {
            zd t = zmk(0.0, 0.0);
            for (int i = jj + 1; i < l; i++) t = zadd(t, zmul(zload(tau, i * l + jj), om[i]));
            om[jj] = zsub(omd[jj], t);
        }
        for (int jj = 0; jj < l - 1; jj++) {
            zd t = zmk(0.0, 0.0);
            for (int i = jj + 1; i < l - 1; i++) t = zadd(t, zmul(zload(tau, i * l + jj), om[i + 1]));
            omdd[jj] = zadd(om[jj + 1], t);
        }
    }
};

// Seed update except the residual (:505-511): x += omega_0 r_0; p_0 += (-omega_{l-1}) p_l;
// for i = 1..l-1: p_0 += (-omega_{i-1}) p_i, x += omega_dd_{i-1} r_i.
__global__ __launch_bounds__(256) void k_zsh_seed(long long n, int l, zd *R, zd *P, zd *x, const Ctl *c) {
    if (ctl_stopped(c)) return;
    const zd *om = (const zd *)&c->s[ZS_OMEGA], *omdd = (const zd *)&c->s[ZS_OMDD];
    ZLOOP(e, n) {
        zd xe = zadd(x[e], zmul(om[0], R[e]));
        zd p0 = zadd(P[e], zmul(zneg(om[l - 1]), P[(size_t)l * n + e]));
        for (int i = 1; i < l; i++) {
            p0 = zadd(p0, zmul(zneg(om[i - 1]), P[(size_t)i * n + e]));
            xe = zadd(xe, zmul(omdd[i - 1], R[(size_t)i * n + e]));
        }
        x[e] = xe;
        P[e] = p0;
    }
}

__device__ double zsh_comb(int n, int r) {         // binomial coefficient (util_shift.c:55-63)
    if (n < 0 || r < 0 || n < r) return 0.0;
    double v = 1.0;
    for (int i = 1; i <= r; i++) v = v * (double)(n - r + i) / (double)i;
    return v;
}

// Per-shift MR scalars (:515-549): Horner re-expansion of omega at sigma_k (zhorner,
// util_shift.c:29-53), e = theta pi, theta *= psi, omega_shift_d via the binomial
// matrix M (:374-380, powers by repeated products), omega_shift_dd via tau.
__global__ void k_zsh_sc2(int ns, int l, ZSh *sh, const double *tau, const Ctl *c) {
    if (ctl_stopped(c)) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    ZSh &z = sh[k];   // in place: a copy of the whole state spills to scratch
    const zd *om = (const zd *)&c->s[ZS_OMEGA];
    const zd sg = z.sigma, msg = zneg(sg);
    zd out[ZS_LMAX + 1];
    out[l] = zneg(om[l - 1]);
    for (int i = l - 1; i > 0; i--) out[i] = zsub(zmul(msg, out[i + 1]), om[i - 1]);
    {
        const zd t = zmul(msg, out[1]);
        out[0] = zmk(t.x + 1.0, t.y);
    }
    for (int i = 1; i < l; i++)
        for (int j = l - 1; j >= i; j--) out[j] = zadd(out[j], zmul(msg, out[j + 1]));
    for (int i = 1; i <= l; i++) out[i] = zdiv(zneg(out[i]), out[0]);
    const zd psi = out[0];
    const zd e = zmul(z.theta, z.pi);
    z.theta = zmul(z.theta, psi);
    zd pw[ZS_LMAX];                                  // sigma^d, d = 0..l-1
    pw[0] = zmk(1.0, 0.0);
    for (int d = 1; d < l; d++) pw[d] = zmul(pw[d - 1], sg);
    zd osd[ZS_LMAX];
    for (int j = 0; j < l; j++) {
        zd acc = zmk(0.0, 0.0);
        for (int i = j; i < l; i++) {
            const double cb = zsh_comb(i, j);
            const zd m = zmk(pw[i - j].x * cb, pw[i - j].y * cb);
            acc = zadd(acc, zmul(m, out[i + 1]));
        }
        osd[j] = acc;
    }
    z.cx[0] = zdiv(osd[0], e);
    for (int j = 0; j < l - 1; j++) {
        zd acc = zmk(0.0, 0.0);
        for (int i = j + 1; i < l - 1; i++) acc = zadd(acc, zmul(zload(tau, i * l + j), osd[i + 1]));
        z.cx[j + 1] = zdiv(zadd(osd[j + 1], acc), e);
    }
    z.cpsi = zdiv(zmk(1.0, 0.0), psi);
}

// Shifted MR update (:551-557): x_k += cx_0 r_0; p_s,0 += (-omega_{l-1}) p_s,l;
// for i = 1..l-1: p_s,0 += (-omega_{i-1}) p_s,i, x_k += cx_i r_i; p_s,0 *= 1/psi.
__global__ __launch_bounds__(256) void k_zsh_v3(long long n, int l, zd *PS, zd *XS, const zd *R, const ZSh *sh,
                                                const Ctl *c) {
    if (ctl_stopped(c)) return;
    const int k = blockIdx.y;
    const int l2 = l + 1;
    const ZSh z = sh[k];
    const zd *om = (const zd *)&c->s[ZS_OMEGA];
    zd *ps = PS + (size_t)k * l2 * n, *xk = XS + (size_t)(k + 1) * n;
    ZLOOP(e, n) {
        zd xe = zadd(xk[e], zmul(z.cx[0], R[e]));
        zd p0 = zadd(ps[e], zmul(zneg(om[l - 1]), ps[(size_t)l * n + e]));
        for (int i = 1; i < l; i++) {
            p0 = zadd(p0, zmul(zneg(om[i - 1]), ps[(size_t)i * n + e]));
            xe = zadd(xe, zmul(z.cx[i], R[(size_t)i * n + e]));
        }
        xk[e] = xe;
        ps[e] = zmul(z.cpsi, p0);
    }
}

// Every per-shift vector operation of one outer iteration in one pass: k_zsh_v1 and
// k_zsh_v2 of each BiCG step j (:426-437, :452-461) and k_zsh_v3 (:551-557).
// p_shift[k][1..l] live only within an outer iteration, so a thread keeps them in
// registers and reads and writes only p_shift[k][0] (at PS + k (l+1) n) and x_k (at
// x + (k+1) n). The seed's r_i before step j's update come from the snapshots k_zsh_r
// wrote, r_j after it from the next step's snapshot (the last step's has its own slot,
// since the MR part's Gram-Schmidt then rewrites r_2..r_l in place), and the MR update
// reads R as k_zsh_v3 does. The same complex operations in the same order as those
// kernels, so the iterates are bitwise theirs; each shift moves 4 n complex values per
// outer iteration. A workgroup row (blockIdx.y) takes ZSH_KCHUNK shifts, their
// coefficients staged in LDS; four shifts' loads are in flight per thread.
#define ZSH_KCHUNK 16
template <int L>
__global__ __launch_bounds__(256) void k_zsh_fused(long long n, int ns, zd *PS, zd *XS, const zd *R,
                                                   const zd *snap, const ZSh *sh, const Ctl *c) {
    if (ctl_stopped(c)) return;
    constexpr int NS = L * (L + 1) / 2;
    constexpr int CW = 6 * L + L + 2;                 // steps, cx, cpsi, sigma
    constexpr int U = 4;
    __shared__ zd cks[ZSH_KCHUNK][CW];
    const int kb = (int)blockIdx.y * ZSH_KCHUNK, ke = min(ns, kb + ZSH_KCHUNK);
    for (int e = threadIdx.x; e < (ke - kb) * CW; e += 256) {
        const ZSh &z = sh[kb + e / CW];
        const int w = e % CW;
        cks[e / CW][w] = w < 6 * L ? z.ck[w / 6][w % 6] : w < 7 * L ? z.cx[w - 6 * L] : (w == 7 * L ? z.cpsi : z.sigma);
    }
    const zd *om = (const zd *)&c->s[ZS_OMEGA];
    zd mom[L];
#pragma unroll
    for (int i = 0; i < L; i++) mom[i] = zneg(om[i]);
    __syncthreads();
    const size_t l2n = (size_t)(L + 1) * n;
    ZLOOP(e, n) {
        zd sn[NS + 1], rf[L];
#pragma unroll
        for (int q = 0; q <= NS; q++) sn[q] = snap[(size_t)q * n + e];
#pragma unroll
        for (int i = 0; i < L; i++) rf[i] = R[(size_t)i * n + e];
        for (int k0 = kb; k0 < ke; k0 += U) {
            zd pl[U], xl[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int k = min(k0 + u, ke - 1);
                pl[u] = PS[(size_t)k * l2n + e];
                xl[u] = XS[(size_t)(k + 1) * n + e];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (k0 + u >= ke) break;
                const zd *cf = cks[k0 + u - kb];
                const zd ms = zneg(cf[7 * L + 1]);
                zd p[L + 1];
                p[0] = pl[u];
                zd xk = xl[u];
#pragma unroll
                for (int j = 0; j < L; j++) {
                    const zd cA = cf[6 * j], cB = cf[6 * j + 1], cC = cf[6 * j + 2], d1 = cf[6 * j + 3],
                             d2 = cf[6 * j + 4], as = cf[6 * j + 5];
                    const int s0 = j * (j + 1) / 2;
#pragma unroll
                    for (int i = 0; i <= j; i++) p[i] = zadd(zmul(cA, p[i]), zmul(cB, sn[s0 + i]));
                    p[j + 1] = zmul(cC, sn[s0 + j]);
                    xk = zadd(xk, zmul(as, p[0]));
                    const zd rjn = (j + 1 < L) ? sn[(j + 1) * (j + 2) / 2 + j] : sn[NS];
                    zd v = zadd(p[j + 1], zmul(d1, rjn));
                    v = zmul(d2, v);
                    p[j + 1] = zadd(v, zmul(ms, p[j]));
                }
                zd xe = zadd(xk, zmul(cf[6 * L], rf[0]));
                zd p0 = zadd(p[0], zmul(mom[L - 1], p[L]));
#pragma unroll
                for (int i = 1; i < L; i++) {
                    p0 = zadd(p0, zmul(mom[i - 1], p[i]));
                    xe = zadd(xe, zmul(cf[6 * L + i], rf[i]));
                }
                const int k = k0 + u;
                XS[(size_t)(k + 1) * n + e] = xe;
                PS[(size_t)k * l2n + e] = zmul(cf[7 * L], p0);
            }
        }
    }
}

// Seed residual (:561-564): r_0 += (-omega_d_{i-1}) r_i, i = 1..l; partial |r_0|^2.
__global__ __launch_bounds__(256) void k_zsh_r0(long long n, int l, zd *R, const Ctl *c, double *part) {
    if (ctl_stopped(c)) return;
    const zd *omd = (const zd *)&c->s[ZS_OMD];
    double v[1] = {0.0};
    ZLOOP(e, n) {
        zd r0 = R[e];
        for (int i = 1; i <= l; i++) r0 = zadd(r0, zmul(zneg(omd[i - 1]), R[(size_t)i * n + e]));
        R[e] = r0;
        v[0] += r0.x * r0.x + r0.y * r0.y;
    }
    block_partial<1, 256>(v, part);
}

extern "C" int bkh_zshbicgstabl(const bkh_zmatrix *A, const double *d_b, int l, const double *sigma, int ns,
                                double *d_x, double tol, int max_iter, bkh_info *info, double *shift_true_res) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !d_b || !d_x || !zbl_ok(l) || ns < 0 || ns > 65535 || (ns > 0 && !sigma)) return BKH_BAD_ARGUMENT;
    if (max_iter <= 0) {                 // no iteration: the reference's error test sees 0
        if (info) { memset(info, 0, sizeof *info); info->status = ST_CONVERGED; }
        return 0;
    }
    const long long n = A->n;
    const int G = vec_grid(n, 256), l2 = l + 1;
    const int Gs = std::max(1, (int)std::min<long long>(G, std::max(1LL, 65536LL / std::max(ns, 1))));
    // the fused per-shift pass unless BKH_ZSH_UNFUSED=1 (the per-step kernels, bitwise the same)
    const char *uf = getenv("BKH_ZSH_UNFUSED");
    const bool fused = ns > 0 && !(uf && uf[0] == '1');
    const size_t nsnap = fused ? (size_t)(l * (l + 1) / 2 + 1) * n : 1;
    void *ws = bkh_workspace(2 * carve_size(sizeof(Ctl)) + carve_size(sizeof(double) * 8 * G) +
                             carve_size(sizeof(zd) * l * l) + carve_size(sizeof(ZSh) * std::max(ns, 1)) +
                             carve_size(sizeof(double) * 2 * std::max(ns, 1)) +
                             (2 * l2 + 2) * carve_size(sizeof(zd) * n) +
                             carve_size(sizeof(zd) * (size_t)n * l2 * std::max(ns, 1)) + carve_size(sizeof(zd) * nsnap));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1), *co = cv.take<Ctl>(1);
    double *part = cv.take<double>(8 * G);
    double *tau = (double *)cv.take<zd>((size_t)l * l);
    ZSh *sh = cv.take<ZSh>(std::max(ns, 1));
    double *dsig = cv.take<double>(2 * std::max(ns, 1));
    zd *R = cv.take<zd>((size_t)l2 * n), *P = cv.take<zd>((size_t)l2 * n), *rs = cv.take<zd>(n), *tmp = cv.take<zd>(n);
    zd *PS = cv.take<zd>((size_t)n * l2 * std::max(ns, 1));
    zd *snap = fused ? cv.take<zd>(nsnap) : nullptr;
    const zd *b = (const zd *)d_b;
    zd *x = (zd *)d_x;
    hipStream_t s = t->stream;
    HIP_OK(hipEventRecord(t->ev0, s));
    if (ctl_init(t, c, tol, max_iter)) return BKH_DEVICE_ERROR;
    if (ns > 0) HIP_OK(hipMemcpyAsync(dsig, sigma, sizeof(double) * 2 * ns, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemsetAsync(P, 0, sizeof(zd) * l2 * n, s));
    HIP_OK(hipMemsetAsync(R, 0, sizeof(zd) * l2 * n, s));
    HIP_OK(hipMemsetAsync(tau, 0, sizeof(zd) * l * l, s));
    HIP_OK(hipMemsetAsync(PS, 0, sizeof(zd) * (size_t)n * l2 * std::max(ns, 1), s));
    hipLaunchKernelGGL(k_zsh_setup, dim3(1), dim3(256), 0, s, c, sh, (const double *)dsig, ns, l);
    launch_zspmv(s, A, false, x, tmp, nullptr);                // :322-329
    hipLaunchKernelGGL(k_zinit, dim3(G), dim3(256), 0, s, n, b, (const zd *)tmp, R, P, rs, (zd *)nullptr, part);
    launch_finalize<3>(s, part, G, c, ZInitFin());
    LAUNCH_OK();
    const int SB = (ns + 63) / 64;
    hipGraphExec_t ge = capture_cached(t, {0x2B18, (uintptr_t)A, (uintptr_t)b, (uintptr_t)x, (uintptr_t)c,
                                           (uintptr_t)part, (uintptr_t)R, (uintptr_t)l, (uintptr_t)ns,
                                           (uintptr_t)PS, (uintptr_t)snap}, s, 1, [=]() {
        for (int j = 0; j < l; j++) {
            launch_zdots<1>(s, n, zdots(rs, R + (size_t)j * n), c, part);
            launch_finalize<2>(s, part, G, c, ZshBetaFin{j, l});
            hipLaunchKernelGGL(k_zsh_p, dim3(G), dim3(256), 0, s, n, j, P, (const zd *)R, (const Ctl *)c);
            launch_zspmv(s, A, false, P + (size_t)j * n, P + (size_t)(j + 1) * n, c);
            launch_zdots<1>(s, n, zdots(rs, P + (size_t)(j + 1) * n), c, part);
            launch_finalize<2>(s, part, G, c, ZshAlphaFin());
            hipLaunchKernelGGL(k_zsh_x, dim3(G), dim3(256), 0, s, n, x, (const zd *)P, (const Ctl *)c);
            if (ns > 0) {
                hipLaunchKernelGGL(k_zsh_sc1, dim3(SB), dim3(64), 0, s, ns, sh, (const Ctl *)c, j);
                if (!fused)
                    hipLaunchKernelGGL(k_zsh_v1, dim3(Gs, ns), dim3(256), 0, s, n, j, l2, PS, x, (const zd *)R,
                                       (const ZSh *)sh, (const Ctl *)c);
            }
            hipLaunchKernelGGL(k_zsh_r, dim3(G), dim3(256), 0, s, n, j, R, (const zd *)P, (const Ctl *)c, snap, l);
            launch_zspmv(s, A, false, R + (size_t)j * n, R + (size_t)(j + 1) * n, c);
            if (ns > 0 && !fused)
                hipLaunchKernelGGL(k_zsh_v2, dim3(Gs, ns), dim3(256), 0, s, n, j, l2, PS, (const zd *)R,
                                   (const ZSh *)sh, (const Ctl *)c);
            hipLaunchKernelGGL(k_zsh_rot, dim3(std::max(SB, 1)), dim3(64), 0, s, ns, sh, c);
        }
        for (int j = 0; j < l; j++) {
            zd *rj = R + (size_t)(j + 1) * n;
            for (int i = 0; i < j; i++) {
                const zd *ri = R + (size_t)(i + 1) * n;
                launch_zdots<1>(s, n, zdots(ri, rj), c, part);
                launch_finalize<2>(s, part, G, c, ZblTauFin{i, j, l, tau});
                hipLaunchKernelGGL(k_zbl_mgs, dim3(G), dim3(256), 0, s, n, rj, ri, (const double *)tau, j * l + i,
                                   (const Ctl *)c);
            }
            launch_zdots<2>(s, n, zdots(rj, rj, rj, R), c, part);
            launch_finalize<4>(s, part, G, c, ZshSigFin{j, l, tau});
        }
        hipLaunchKernelGGL(k_zsh_seed, dim3(G), dim3(256), 0, s, n, l, R, P, x, (const Ctl *)c);
        if (ns > 0) {
            hipLaunchKernelGGL(k_zsh_sc2, dim3(SB), dim3(64), 0, s, ns, l, sh, (const double *)tau, (const Ctl *)c);
            if (fused) {
                const dim3 gf((unsigned)vec_grid(n, 256, 1), (unsigned)((ns + ZSH_KCHUNK - 1) / ZSH_KCHUNK));
                switch (l) {
#define ZCASE(L_) case L_: hipLaunchKernelGGL((k_zsh_fused<L_>), gf, dim3(256), 0, s, n, ns, PS, x, (const zd *)R, \
                                              (const zd *)snap, (const ZSh *)sh, (const Ctl *)c); break;
                    ZCASE(1) ZCASE(2) ZCASE(3) ZCASE(4) ZCASE(5) ZCASE(6) ZCASE(7) ZCASE(8)
#undef ZCASE
                }
            } else {
                hipLaunchKernelGGL(k_zsh_v3, dim3(Gs, ns), dim3(256), 0, s, n, l, PS, x, (const zd *)R,
                                   (const ZSh *)sh, (const Ctl *)c);
            }
        }
        hipLaunchKernelGGL(k_zsh_r0, dim3(G), dim3(256), 0, s, n, l, R, (const Ctl *)c, part);
        launch_finalize<1>(s, part, G, c, ZCheckFin());
    });
    if (!ge) { bkh_set_error("hipGraph capture of the complex shifted BiCGSTAB(l) iteration failed"); return BKH_DEVICE_ERROR; }
    if (run_graph_until_done(t, ge, c, max_iter + 2)) return BKH_DEVICE_ERROR;
    int ret = zsolve_end(t, A, b, x, c, co, tmp, part, info);
    if (ret < -4) return ret;
    if (shift_true_res) {
        if (ztrue_res(t, A, b, x, zmk(0.0, 0.0), tmp, co, part, &shift_true_res[0])) return BKH_DEVICE_ERROR;
        for (int k = 0; k < ns; k++)
            if (ztrue_res(t, A, b, x + (size_t)(k + 1) * n, zmk(sigma[2 * k], sigma[2 * k + 1]), tmp, co, part,
                          &shift_true_res[k + 1]))
                return BKH_DEVICE_ERROR;
    }
    return ret;
}

// Average duration of `reps` back-to-back complex SpMV launches (HIP events on the
// stream they run on), x = 1; for the roofline of the complex product.
extern "C" int bkh_zspmv_bench(const bkh_zmatrix *A, char op, int reps, double *avg_ms) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !avg_ms || reps < 1 || (op != 'n' && op != 't')) return BKH_BAD_ARGUMENT;
    const long long n = A->n;
    void *ws = bkh_workspace(2 * carve_size(sizeof(zd) * n));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    zd *x = cv.take<zd>(n), *y = cv.take<zd>(n);
    hipStream_t s = t->stream;
    HIP_OK(hipMemsetAsync(x, 0, sizeof(zd) * n, s));
    launch_zspmv(s, A, op == 't', x, y, nullptr, false);
    HIP_OK(hipEventRecord(t->ev0, s));
    for (int k = 0; k < reps; k++) launch_zspmv(s, A, op == 't', x, y, nullptr, false);
    HIP_OK(hipEventRecord(t->ev1, s));
    LAUNCH_OK();
    HIP_OK(hipEventSynchronize(t->ev1));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    *avg_ms = ms / reps;
    return 0;
}

// ---- block BiCGSTAB(l) (blbicgstabl.c:255-486) and its thin QR (zqreco, myqr.c:41-76)
//
// Blocks are m x s column-major (ld m), as zmat. The reference's Gram products are
// bilinear (cblas_zgemm CblasTrans: Rs^T V); the QR's Gram is V^H V. zqreco is
// CholeskyQR2 here: R has a real positive diagonal where zgeqrf's R has a real
// diagonal of either sign, so the two differ by Q -> Q D, R -> D R with D = diag(+-1),
// under which the block iterates are invariant (as for the real block solvers). The
// s x s algebra (zgesv, zgetrf/zgetri, products, Cholesky) runs in one workgroup.
#define ZB_SMAX 16
#define ZB_SN (ZB_SMAX * ZB_SMAX)

struct ZBState {
    zd alpha[ZB_SN], beta[ZB_SN], t[ZB_SN], xi[ZB_SN], g0[ZB_SN], g1[ZB_SN];
    zd R1[ZB_SN], R1i[ZB_SN], R2i[ZB_SN], work[2 * ZB_SN];
    double fnb;
};

// C = A B (s x s, column-major, LDS or global), one thread per entry.
__device__ void zsm_mul(int s, const zd *A, const zd *B, zd *C) {
    const int t = threadIdx.x;
    if (t < s * s) {
        const int i = t % s, j = t / s;
        zd acc = zmk(0.0, 0.0);
        for (int k = 0; k < s; k++) acc = zadd(acc, zmul(A[i + k * s], B[k + j * s]));
        C[t] = acc;
    }
    __syncthreads();
}

__device__ void zsm_copy(int s, const zd *A, zd *B) {
    if ((int)threadIdx.x < s * s) B[threadIdx.x] = A[threadIdx.x];
    __syncthreads();
}

// X = A^-1 B (Gauss-Jordan with partial pivoting on [A | B], max |re|+|im| pivot);
// work: 2 s^2 entries, piv: 2 ints in LDS. Returns false (uniformly) if singular.
__device__ bool zsm_solve(int s, const zd *A, const zd *B, zd *X, zd *work, int *piv) {
    const int t = threadIdx.x;
    zd *M = work;                                  // s x 2s, column-major, ld s
    if (t < s * s) { M[t] = A[t]; M[s * s + t] = B[t]; }
    if (t == 0) piv[1] = 1;
    __syncthreads();
    for (int k = 0; k < s; k++) {
        if (t == 0) {
            int p = k;
            double best = fabs(M[k + k * s].x) + fabs(M[k + k * s].y);
            for (int i = k + 1; i < s; i++) {
                const double a = fabs(M[i + k * s].x) + fabs(M[i + k * s].y);
                if (a > best) { best = a; p = i; }
            }
            piv[0] = p;
            if (!(best > 0.0)) piv[1] = 0;
        }
        __syncthreads();
        const int p = piv[0];
        if (p != k && t < 2 * s) {
            const zd a = M[k + t * s], b = M[p + t * s];
            M[k + t * s] = b;
            M[p + t * s] = a;
        }
        __syncthreads();
        const zd inv = piv[1] ? zdiv(zmk(1.0, 0.0), M[k + k * s]) : zmk(0.0, 0.0);
        __syncthreads();
        if (t < 2 * s) M[k + t * s] = zmul(M[k + t * s], inv);
        __syncthreads();
        for (int e = t; e < s * 2 * s; e += blockDim.x) {
            const int i = e % s, c = e / s;
            if (i != k && c != k) M[i + c * s] = zsub(M[i + c * s], zmul(M[i + k * s], M[k + c * s]));
        }
        __syncthreads();
        if (t < s && t != k) M[t + k * s] = zmk(0.0, 0.0);
        __syncthreads();
    }
    if (t < s * s) X[t] = M[s * s + t];
    __syncthreads();
    return piv[1] != 0;
}

// Upper Cholesky factor of a Hermitian positive definite G (G = R^H R).
__device__ bool zsm_chol(int s, const zd *G, zd *R, int *flag) {
    const int t = threadIdx.x;
    if (t == 0) *flag = 1;
    if (t < s * s) {
        const int i = t % s, j = t / s;
        R[t] = (i <= j) ? G[t] : zmk(0.0, 0.0);
    }
    __syncthreads();
    for (int k = 0; k < s; k++) {
        if (t == 0) {
            const double d = R[k + k * s].x;
            if (!(d > 0.0)) *flag = 0;
            R[k + k * s] = zmk(d > 0.0 ? sqrt(d) : 1.0, 0.0);
        }
        __syncthreads();
        const double rkk = R[k + k * s].x;
        if (t > k && t < s) R[k + t * s] = zmk(R[k + t * s].x / rkk, R[k + t * s].y / rkk);
        __syncthreads();
        if (t < s * s) {
            const int i = t % s, j = t / s;
            if (i > k && j >= i) R[i + j * s] = zsub(R[i + j * s], zmul(zconj(R[k + i * s]), R[k + j * s]));
        }
        __syncthreads();
    }
    return *flag != 0;
}

// Inverse of an upper-triangular R (column j by back substitution, thread j).
__device__ void zsm_triu_inv(int s, const zd *R, zd *Ri) {
    const int t = threadIdx.x;
    if (t < s * s) Ri[t] = zmk(0.0, 0.0);
    __syncthreads();
    if (t < s) {
        const int j = t;
        for (int i = j; i >= 0; i--) {
            zd acc = (i == j) ? zmk(1.0, 0.0) : zmk(0.0, 0.0);
            for (int k = i + 1; k <= j; k++) acc = zsub(acc, zmul(R[i + k * s], Ri[k + j * s]));
            Ri[i + j * s] = zdiv(acc, R[i + i * s]);
        }
    }
    __syncthreads();
}

// Partial Gram of one workgroup's rows: G[i][j] = sum_r f(X[r,i]) Y[r,j], f = conj when
// CONJ; 64-row tiles staged in LDS, one thread per entry.
template <bool CONJ>
__global__ __launch_bounds__(256) void k_zgram(long long m, int s, const zd *X, const zd *Y, const Ctl *c,
                                               double *part) {
    if (c && ctl_stopped(c)) return;
    __shared__ zd tx[64 * ZB_SMAX], ty[64 * ZB_SMAX];
    const int t = threadIdx.x, ss = s * s;
    const int ei = t % s, ej = t / s;
    zd acc = zmk(0.0, 0.0);
    for (long long r0 = (long long)blockIdx.x * 64; r0 < m; r0 += (long long)gridDim.x * 64) {
        const int rows = (int)min(64LL, m - r0);
        for (int e = t; e < 64 * s; e += 256) {
            const int r = e % 64, q = e / 64;
            zd a = zmk(0.0, 0.0), b = zmk(0.0, 0.0);
            if (r < rows) { a = X[(size_t)q * m + r0 + r]; b = Y[(size_t)q * m + r0 + r]; }
            tx[q * 64 + r] = CONJ ? zconj(a) : a;
            ty[q * 64 + r] = b;
        }
        __syncthreads();
        if (t < ss)
            for (int r = 0; r < rows; r++) acc = zadd(acc, zmul(tx[ei * 64 + r], ty[ej * 64 + r]));
        __syncthreads();
    }
    if (t < ss) {
        part[((size_t)blockIdx.x * ss + t) * 2] = acc.x;
        part[((size_t)blockIdx.x * ss + t) * 2 + 1] = acc.y;
    }
}

static inline int zgram_grid(long long m) { return (int)std::max(1LL, std::min((m + 63) / 64, 256LL)); }

// The same partial Gram for s = S <= 4: every thread accumulates all S x S entries over
// its rows (grid-stride) in registers, then the workgroup sums them (wave sums, then the
// four waves in order) into the block's partial row.
template <int S, bool CONJ>
__global__ __launch_bounds__(256) void k_zgram_r(long long m, const zd *X, const zd *Y, const Ctl *c, double *part) {
    if (c && ctl_stopped(c)) return;
    constexpr int SS = S * S;
    double re[SS], im[SS];
#pragma unroll
    for (int e = 0; e < SS; e++) { re[e] = 0.0; im[e] = 0.0; }
    ZLOOP(r, m) {
        zd x[S], y[S];
#pragma unroll
        for (int i = 0; i < S; i++) {
            x[i] = X[(size_t)i * m + r];
            if (CONJ) x[i] = zconj(x[i]);
            y[i] = Y[(size_t)i * m + r];
        }
#pragma unroll
        for (int j = 0; j < S; j++)
#pragma unroll
            for (int i = 0; i < S; i++) {
                const zd pr = zmul(x[i], y[j]);
                re[i + j * S] += pr.x;
                im[i + j * S] += pr.y;
            }
    }
    __shared__ double sh[2 * SS * 4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int e = 0; e < SS; e++) {
        re[e] = wave_sum(re[e]);
        im[e] = wave_sum(im[e]);
    }
    if (lane == 0) {
#pragma unroll
        for (int e = 0; e < SS; e++) { sh[(2 * e) * 4 + w] = re[e]; sh[(2 * e + 1) * 4 + w] = im[e]; }
    }
    __syncthreads();
    if ((int)threadIdx.x < 2 * SS) {
        const double *v = sh + threadIdx.x * 4;
        part[(size_t)blockIdx.x * 2 * SS + threadIdx.x] = ((v[0] + v[1]) + v[2]) + v[3];
    }
}

// Partial Gram rows (zgram_grid(m) of them) of f(X)^T Y, f = conj when CONJ.
template <bool CONJ>
static void zgram_launch(hipStream_t st, long long m, int s, const zd *X, const zd *Y, const Ctl *c, double *part) {
    const int Gg = zgram_grid(m);
    switch (s) {
    case 1: hipLaunchKernelGGL((k_zgram_r<1, CONJ>), dim3(Gg), dim3(256), 0, st, m, X, Y, c, part); break;
    case 2: hipLaunchKernelGGL((k_zgram_r<2, CONJ>), dim3(Gg), dim3(256), 0, st, m, X, Y, c, part); break;
    case 3: hipLaunchKernelGGL((k_zgram_r<3, CONJ>), dim3(Gg), dim3(256), 0, st, m, X, Y, c, part); break;
    case 4: hipLaunchKernelGGL((k_zgram_r<4, CONJ>), dim3(Gg), dim3(256), 0, st, m, X, Y, c, part); break;
    default: hipLaunchKernelGGL((k_zgram<CONJ>), dim3(Gg), dim3(256), 0, st, m, s, X, Y, c, part);
    }
}

enum { ZB_QR1 = 0, ZB_QR2_INIT = 1, ZB_QR2_ITER = 2, ZB_BETA = 3, ZB_ALPHA = 4, ZB_XI = 5 };

// Sums the Gram partials (fixed order) and runs the s x s step `op`:
//   ZB_BETA (j): on j == 0 first g0 = (-omega[l-1]) g0 (:339-340); g1 = Rs^T V_j;
//               beta = alpha (g0^-1 g1); g0 = g1                          (:344-350)
//   ZB_ALPHA:   alpha = (Rs^T P_{j+1})^-1 g1                              (:363-366)
//   ZB_QR1:     R1 = chol(V^H V), R1i = R1^-1
//   ZB_QR2_*:   R2 = chol(V^H V), R2i = R2^-1; R = R2 R1 (zqreco's R);
//               INIT: xi = R (:327); ITER: t = R (:373)
//   ZB_XI:      xi = t xi, t = t^-1 (zgetrf/zgetri of the triangular t, :376-380);
//               no partials
__global__ __launch_bounds__(256) void k_zblk_fin(const double *part, int nblocks, int s, int op, int j, int l,
                                                  ZBState *st, Ctl *c) {
    if (ctl_stopped(c)) return;
    __shared__ zd G[ZB_SN], T1[ZB_SN], T2[ZB_SN], W[2 * ZB_SN];
    __shared__ int flag[2];
    const int t = threadIdx.x, ss = s * s;
    if (op == ZB_XI) {
        zsm_mul(s, st->t, st->xi, G);
        zsm_copy(s, G, st->xi);
        zsm_copy(s, st->t, T1);
        zsm_triu_inv(s, T1, st->t);
        return;
    }
    {
        // the partial rows in 256 / ss interleaved groups, then the groups in order
        const int ng = 256 / ss, e = t % ss, g = t / ss;
        double re = 0.0, im = 0.0;
        if (g < ng)
            for (int b = g; b < nblocks; b += ng) {
                re += part[((size_t)b * ss + e) * 2];
                im += part[((size_t)b * ss + e) * 2 + 1];
            }
        double *W2 = (double *)W;
        __syncthreads();
        if (g < ng) { W2[2 * (g * ss + e)] = re; W2[2 * (g * ss + e) + 1] = im; }
        __syncthreads();
        if (t < ss) {
            double sr = 0.0, si = 0.0;
            for (int k = 0; k < ng; k++) { sr += W2[2 * (k * ss + t)]; si += W2[2 * (k * ss + t) + 1]; }
            G[t] = zmk(sr, si);
        }
    }
    __syncthreads();
    if (op == ZB_BETA) {
        if (j == 0 && t < ss) st->g0[t] = zmul(zneg(zload(c->s, ZS_OMEGA / 2 + l - 1)), st->g0[t]);
        __syncthreads();
        zsm_copy(s, G, st->g1);
        if (!zsm_solve(s, st->g0, G, T1, W, flag)) { if (t == 0) ctl_breakdown(c); return; }
        zsm_mul(s, st->alpha, T1, st->beta);
        zsm_copy(s, G, st->g0);
        return;
    }
    if (op == ZB_ALPHA) {
        if (!zsm_solve(s, G, st->g1, T1, W, flag)) { if (t == 0) ctl_breakdown(c); return; }
        zsm_copy(s, T1, st->alpha);
        return;
    }
    if (!zsm_chol(s, G, T1, flag)) { if (t == 0) ctl_breakdown(c); return; }
    if (op == ZB_QR1) {
        zsm_copy(s, T1, st->R1);
        zsm_triu_inv(s, T1, st->R1i);
        return;
    }
    zsm_triu_inv(s, T1, st->R2i);
    zsm_mul(s, T1, st->R1, T2);                    // R = R2 R1
    zsm_copy(s, T2, op == ZB_QR2_INIT ? st->xi : st->t);
}

// k_zpan / k_zpan_vt for s = S known at compile time: the row's entries stay in
// registers and the s x s matrices in LDS (the runtime-s forms index private arrays and
// spill them); the same operations in the same order.
template <int S>
__global__ __launch_bounds__(256) void k_zpan_t(long long m, const zd *X, const zd *M, zd a, const zd *Y, zd b,
                                                zd *O, const Ctl *c) {
    if (c && ctl_stopped(c)) return;
    __shared__ zd Ms[S * S];
    for (int e = threadIdx.x; e < S * S; e += 256) Ms[e] = M[e];
    __syncthreads();
    ZLOOP(r, m) {
        zd xr[S];
#pragma unroll
        for (int l = 0; l < S; l++) xr[l] = X[(size_t)l * m + r];
#pragma unroll
        for (int q = 0; q < S; q++) {
            zd acc = zmk(0.0, 0.0);
#pragma unroll
            for (int l = 0; l < S; l++) acc = zadd(acc, zmul(xr[l], Ms[l + q * S]));
            zd o = zmul(a, acc);
            if (Y) o = zadd(o, zmul(b, Y[(size_t)q * m + r]));
            O[(size_t)q * m + r] = o;
        }
    }
}

template <int S>
__global__ __launch_bounds__(256) void k_zpan_vt_t(long long m, zd *V, const zd *P1, const ZBState *st,
                                                   const Ctl *c) {
    if (ctl_stopped(c)) return;
    __shared__ zd al[S * S], tt[S * S];
    for (int e = threadIdx.x; e < S * S; e += 256) {
        const int l = e % S, q = e / S;
        al[e] = st->alpha[l + q * S];
        tt[e] = st->t[l + q * S];
    }
    __syncthreads();
    ZLOOP(r, m) {
        zd v[S], p[S], w[S];
#pragma unroll
        for (int l = 0; l < S; l++) { p[l] = P1[(size_t)l * m + r]; v[l] = V[(size_t)l * m + r]; }
#pragma unroll
        for (int q = 0; q < S; q++) {
            zd acc = zmk(0.0, 0.0);
#pragma unroll
            for (int l = 0; l < S; l++) acc = zadd(acc, zmul(p[l], al[l + q * S]));
            w[q] = zadd(zmul(zmk(-1.0, 0.0), acc), zmul(zmk(1.0, 0.0), v[q]));
        }
#pragma unroll
        for (int q = 0; q < S; q++) {
            zd acc = zmk(0.0, 0.0);
#pragma unroll
            for (int l = 0; l < S; l++) acc = zadd(acc, zmul(w[l], tt[l + q * S]));
            V[(size_t)q * m + r] = zmul(zmk(1.0, 0.0), acc);
        }
    }
}

#define ZB_SWITCH(s_, CALL)                                                                   \
    switch (s_) {                                                                             \
    case 1: CALL(1); break; case 2: CALL(2); break; case 3: CALL(3); break; case 4: CALL(4); break; \
    case 5: CALL(5); break; case 6: CALL(6); break; case 7: CALL(7); break; case 8: CALL(8); break; \
    case 9: CALL(9); break; case 10: CALL(10); break; case 11: CALL(11); break; case 12: CALL(12); break; \
    case 13: CALL(13); break; case 14: CALL(14); break; case 15: CALL(15); break; default: CALL(16); break; \
    }

static void zpan_launch(hipStream_t st, int G, long long m, int s, const zd *X, const zd *M, zd a, const zd *Y, zd b,
                        zd *O, const Ctl *c) {
#define ZPAN_CALL(S_) hipLaunchKernelGGL((k_zpan_t<S_>), dim3(G), dim3(256), 0, st, m, X, M, a, Y, b, O, c)
    ZB_SWITCH(s, ZPAN_CALL)
#undef ZPAN_CALL
}

static void zpan_vt_launch(hipStream_t st, int G, long long m, int s, zd *V, const zd *P1, const ZBState *bs,
                           const Ctl *c) {
#define ZPANVT_CALL(S_) hipLaunchKernelGGL((k_zpan_vt_t<S_>), dim3(G), dim3(256), 0, st, m, V, P1, bs, c)
    ZB_SWITCH(s, ZPANVT_CALL)
#undef ZPANVT_CALL
}

// One row per thread, all s columns: O[r,:] = a (X[r,:] M) (+ b Y[r,:]); M from the
// state (s x s). X may alias O only when Y is null or Y == O (row-local).
__global__ __launch_bounds__(256) void k_zpan(long long m, int s, const zd *X, const zd *M, zd a, const zd *Y, zd b,
                                              zd *O, const Ctl *c) {
    if (c && ctl_stopped(c)) return;
    ZLOOP(r, m) {
        zd xr[ZB_SMAX];
        for (int l = 0; l < s; l++) xr[l] = X[(size_t)l * m + r];
        for (int q = 0; q < s; q++) {
            zd acc = zmk(0.0, 0.0);
            for (int l = 0; l < s; l++) acc = zadd(acc, zmul(xr[l], M[l + q * s]));
            zd o = zmul(a, acc);
            if (Y) o = zadd(o, zmul(b, Y[(size_t)q * m + r]));
            O[(size_t)q * m + r] = o;
        }
    }
}

// V_i = ((-1) (P_{i+1} alpha) + 1 V_i) t, row-local (:392-396).
__global__ __launch_bounds__(256) void k_zpan_vt(long long m, int s, zd *V, const zd *P1, const ZBState *st,
                                                 const Ctl *c) {
    if (ctl_stopped(c)) return;
    ZLOOP(r, m) {
        zd v[ZB_SMAX], p[ZB_SMAX];
        for (int l = 0; l < s; l++) { p[l] = P1[(size_t)l * m + r]; v[l] = V[(size_t)l * m + r]; }
        zd w[ZB_SMAX];
        for (int q = 0; q < s; q++) {
            zd acc = zmk(0.0, 0.0);
            for (int l = 0; l < s; l++) acc = zadd(acc, zmul(p[l], st->alpha[l + q * s]));
            w[q] = zadd(zmul(zmk(-1.0, 0.0), acc), zmul(zmk(1.0, 0.0), v[q]));
        }
        for (int q = 0; q < s; q++) {
            zd acc = zmk(0.0, 0.0);
            for (int l = 0; l < s; l++) acc = zadd(acc, zmul(w[l], st->t[l + q * s]));
            V[(size_t)q * m + r] = zmul(zmk(1.0, 0.0), acc);
        }
    }
}

// tmp = sum_i omega_i V_i (from 0: zdscal(0) then zaxpy); P_0 += (-omega_i) P_{i+1},
// i = 0..l-1 (:442-447), over all m s entries.
__global__ __launch_bounds__(256) void k_zblk_upd(long long ms, int l, const zd *V, zd *P, zd *tmp, const Ctl *c) {
    if (ctl_stopped(c)) return;
    const zd *om = (const zd *)&c->s[ZS_OMEGA];
    ZLOOP(e, ms) {
        zd tv = zmk(0.0, 0.0), p0 = P[e];
        for (int i = 0; i < l; i++) {
            tv = zadd(tv, zmul(om[i], V[(size_t)i * ms + e]));
            p0 = zadd(p0, zmul(zneg(om[i]), P[(size_t)(i + 1) * ms + e]));
        }
        tmp[e] = tv;
        P[e] = p0;
    }
}

// error = ||xi||_F / ||B||_F (:454-457): the block solvers' stop test, NaN stops.
__global__ void k_zblk_check(int s, const ZBState *st, Ctl *c) {
    if (ctl_stopped(c)) return;
    if (threadIdx.x != 0) return;
    double acc = 0.0;
    for (int e = 0; e < s * s; e++) acc += st->xi[e].x * st->xi[e].x + st->xi[e].y * st->xi[e].y;
    ctl_check(c, acc);
}

// MR part of the block solver (:413-432): as the shifted solver's (omega_dd unused).
struct ZbSigFin {
    int j, l;
    const double *tau;
    __device__ void operator()(Ctl *c, const double *tot) const {
        ZshSigFin{j, l, tau}(c, tot);
    }
};

__global__ void k_zblk_init(Ctl *c, ZBState *st, int s, int l) {
    const int t = threadIdx.x;
    if (t < s * s) {
        st->g0[t] = (t % s == t / s) ? zmk(1.0, 0.0) : zmk(0.0, 0.0);
        st->alpha[t] = zmk(0.0, 0.0);
    }
    if (t == 0) zstore(c->s, ZS_OMEGA / 2 + l - 1, zmk(1.0, 0.0));
}

// |B|_F^2 -> Ctl::fnb.
__global__ __launch_bounds__(256) void k_znorm2(long long n, const zd *x, double *part) {
    double v[1] = {0.0};
    ZLOOP(i, n) v[0] += x[i].x * x[i].x + x[i].y * x[i].y;
    block_partial<1, 256>(v, part);
}

struct ZFnbFin {
    __device__ void operator()(Ctl *c, const double *tot) const {
        c->fnb = sqrt(tot[0]);
        if (!(c->fnb > 0.0)) ctl_breakdown(c);
    }
};

static void launch_zspmm1(hipStream_t st, const bkh_zmatrix *A, int s, zd alpha, const zd *X, zd beta, zd *Y) {
    zspmm1_launch<false>(st, A->A, A->n, s, alpha, beta, X, Y);
    launch_zsp_tail<true>(st, A->A, false, A->n, s, alpha, X, Y, nullptr);
}

// CholeskyQR2 of V (m x s) in place; op2 = ZB_QR2_INIT or ZB_QR2_ITER.
static void zqr2(hipStream_t st, long long m, int s, zd *V, ZBState *bs, Ctl *c, double *part, int op2) {
    const int Gg = zgram_grid(m), Gp = vec_grid(m, 256);
    zgram_launch<true>(st, m, s, (const zd *)V, (const zd *)V,
                       (const Ctl *)c, part);
    hipLaunchKernelGGL(k_zblk_fin, dim3(1), dim3(256), 0, st, (const double *)part, Gg, s, (int)ZB_QR1, 0, 0, bs, c);
    zpan_launch(st, Gp, m, s, (const zd *)V, (const zd *)bs->R1i, zmk(1.0, 0.0),
                       (const zd *)nullptr, zmk(0.0, 0.0), V, (const Ctl *)c);
    zgram_launch<true>(st, m, s, (const zd *)V, (const zd *)V,
                       (const Ctl *)c, part);
    hipLaunchKernelGGL(k_zblk_fin, dim3(1), dim3(256), 0, st, (const double *)part, Gg, s, op2, 0, 0, bs, c);
    zpan_launch(st, Gp, m, s, (const zd *)V, (const zd *)bs->R2i, zmk(1.0, 0.0),
                       (const zd *)nullptr, zmk(0.0, 0.0), V, (const Ctl *)c);
}

extern "C" int bkh_zblbicgstabl(const bkh_zmatrix *A, int s, const double *d_B, int l, double *d_X, double tol,
                                int max_iter, bkh_info *info) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !d_B || !d_X || !zbl_ok(l) || s < 1 || s > ZB_SMAX) return BKH_BAD_ARGUMENT;
    if (max_iter <= 0) {
        if (info) { memset(info, 0, sizeof *info); info->status = ST_CONVERGED; }
        return 0;
    }
    const long long m = A->n, ms = m * s;
    const int Gv = vec_grid(ms, 256), Gg = zgram_grid(m), Gp = vec_grid(m, 256), l2 = l + 1;
    const int nparts = std::max(Gg * ZB_SN * 2, 8 * Gv);
    void *ws = bkh_workspace(2 * carve_size(sizeof(Ctl)) + carve_size(sizeof(ZBState)) +
                             carve_size(sizeof(double) * nparts) + carve_size(sizeof(zd) * l * l) +
                             (3 * l2 + 2) * carve_size(sizeof(zd) * ms));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1), *co = cv.take<Ctl>(1);
    ZBState *bs = cv.take<ZBState>(1);
    double *part = cv.take<double>(nparts);
    double *tau = (double *)cv.take<zd>((size_t)l * l);
    zd *Pb = cv.take<zd>((size_t)l2 * ms), *Vb = cv.take<zd>((size_t)l2 * ms), *Rb = cv.take<zd>((size_t)l2 * ms);
    zd *Rs = cv.take<zd>(ms), *tmp = cv.take<zd>(ms);
    const zd *B = (const zd *)d_B;
    zd *X = (zd *)d_X;
    auto Pi = [=](int i) { return Pb + (size_t)i * ms; };
    auto Vi = [=](int i) { return Vb + (size_t)i * ms; };
    auto Ri = [=](int i) { return Rb + (size_t)i * ms; };
    const zd one = zmk(1.0, 0.0), mone = zmk(-1.0, 0.0), zero = zmk(0.0, 0.0);
    hipStream_t st = t->stream;
    HIP_OK(hipEventRecord(t->ev0, st));
    if (ctl_init(t, c, tol, max_iter)) return BKH_DEVICE_ERROR;
    HIP_OK(hipMemsetAsync(Pb, 0, sizeof(zd) * l2 * ms, st));
    HIP_OK(hipMemsetAsync(Vb, 0, sizeof(zd) * l2 * ms, st));
    HIP_OK(hipMemsetAsync(tau, 0, sizeof(zd) * l * l, st));
    hipLaunchKernelGGL(k_zblk_init, dim3(1), dim3(256), 0, st, c, bs, s, l);
    // V0 = B - A X0 (:322-326: zcsrmm with -1, zdscal(-1), zaxpy(one, B))
    HIP_OK(hipMemsetAsync(tmp, 0, sizeof(zd) * ms, st));
    launch_zspmm1(st, A, s, mone, X, zero, tmp);
    hipLaunchKernelGGL(k_zinit, dim3(Gv), dim3(256), 0, st, ms, B, (const zd *)tmp, Vi(0), (zd *)nullptr,
                       (zd *)nullptr, (zd *)nullptr, part);
    hipLaunchKernelGGL(k_znorm2, dim3(Gv), dim3(256), 0, st, ms, B, part);
    launch_finalize<1>(st, part, Gv, c, ZFnbFin());
    zqr2(st, m, s, Vi(0), bs, c, part, ZB_QR2_INIT);           // zqreco(V0, xi) (:327)
    HIP_OK(hipMemcpyAsync(Pi(0), Vi(0), sizeof(zd) * ms, hipMemcpyDeviceToDevice, st));
    HIP_OK(hipMemcpyAsync(Rs, Vi(0), sizeof(zd) * ms, hipMemcpyDeviceToDevice, st));
    LAUNCH_OK();
    hipGraphExec_t ge = capture_cached(t, {0x2B19, (uintptr_t)A, (uintptr_t)B, (uintptr_t)X, (uintptr_t)c,
                                           (uintptr_t)part, (uintptr_t)Pb, (uintptr_t)l, (uintptr_t)s}, st, 1, [=]() {
        for (int j = 0; j < l; j++) {
            zgram_launch<false>(st, m, s, (const zd *)Rs, (const zd *)Vi(j),
                               (const Ctl *)c, part);
            hipLaunchKernelGGL(k_zblk_fin, dim3(1), dim3(256), 0, st, (const double *)part, Gg, s, (int)ZB_BETA, j, l,
                               bs, c);
            for (int i = 0; i <= j; i++)                         // P_i = (-1) P_i beta + 1 V_i (:353-358)
                zpan_launch(st, Gp, m, s, (const zd *)Pi(i), (const zd *)bs->beta,
                                   mone, (const zd *)Vi(i), one, Pi(i), (const Ctl *)c);
            launch_zspmm1(st, A, s, one, Pi(j), zero, Pi(j + 1));
            zgram_launch<false>(st, m, s, (const zd *)Rs,
                               (const zd *)Pi(j + 1), (const Ctl *)c, part);
            hipLaunchKernelGGL(k_zblk_fin, dim3(1), dim3(256), 0, st, (const double *)part, Gg, s, (int)ZB_ALPHA, j,
                               l, bs, c);
            // i = 0 (:371-381): tmp = P0 alpha; V0 = V0 - A tmp; qr(V0) -> t; X += tmp xi;
            // xi = t xi; t = t^-1
            zpan_launch(st, Gp, m, s, (const zd *)Pi(0), (const zd *)bs->alpha, one,
                               (const zd *)nullptr, zero, tmp, (const Ctl *)c);
            launch_zspmm1(st, A, s, mone, tmp, one, Vi(0));
            zqr2(st, m, s, Vi(0), bs, c, part, ZB_QR2_ITER);
            zpan_launch(st, Gp, m, s, (const zd *)tmp, (const zd *)bs->xi, one,
                               (const zd *)X, one, X, (const Ctl *)c);
            hipLaunchKernelGGL(k_zblk_fin, dim3(1), dim3(256), 0, st, (const double *)part, 0, s, (int)ZB_XI, j,
                               l, bs, c);
            for (int i = 1; i <= j; i++)                         // :392-396
                zpan_vt_launch(st, Gp, m, s, Vi(i), (const zd *)Pi(i + 1),
                                   (const ZBState *)bs, (const Ctl *)c);
            launch_zspmm1(st, A, s, one, Vi(j), zero, Vi(j + 1));
        }
        for (int i = 0; i <= l; i++)                             // R_i = V_i xi (:406-409)
            zpan_launch(st, Gp, m, s, (const zd *)Vi(i), (const zd *)bs->xi, one,
                               (const zd *)nullptr, zero, Ri(i), (const Ctl *)c);
        for (int j = 0; j < l; j++) {                            // MR (:413-432)
            zd *rj = Ri(j + 1);
            for (int i = 0; i < j; i++) {
                const zd *ri = Ri(i + 1);
                launch_zdots<1>(st, ms, zdots(ri, rj), c, part);
                launch_finalize<2>(st, part, Gv, c, ZblTauFin{i, j, l, tau});
                hipLaunchKernelGGL(k_zbl_mgs, dim3(Gv), dim3(256), 0, st, ms, rj, ri, (const double *)tau, j * l + i,
                                   (const Ctl *)c);
            }
            launch_zdots<2>(st, ms, zdots(rj, rj, rj, Ri(0)), c, part);
            launch_finalize<4>(st, part, Gv, c, ZbSigFin{j, l, tau});
        }
        hipLaunchKernelGGL(k_zblk_upd, dim3(Gv), dim3(256), 0, st, ms, l, (const zd *)Vb, Pb, tmp, (const Ctl *)c);
        launch_zspmm1(st, A, s, mone, tmp, one, Vi(0));
        zpan_launch(st, Gp, m, s, (const zd *)tmp, (const zd *)bs->xi, one,
                           (const zd *)X, one, X, (const Ctl *)c);
        hipLaunchKernelGGL(k_zblk_check, dim3(1), dim3(64), 0, st, s, (const ZBState *)bs, c);
    });
    if (!ge) { bkh_set_error("hipGraph capture of the complex block BiCGSTAB(l) iteration failed"); return BKH_DEVICE_ERROR; }
    if (run_graph_until_done(t, ge, c, max_iter + 2)) return BKH_DEVICE_ERROR;
    HIP_OK(hipEventRecord(t->ev1, st));
    if (ctl_fetch(t, c)) return BKH_DEVICE_ERROR;
    const Ctl hc = *t->ctl_host;
    HIP_OK(hipEventSynchronize(t->ev1));
    bkh_info inf;
    memset(&inf, 0, sizeof inf);
    const int ret = zresult(t, &hc, &inf);
    if (info) {
        // Frobenius true residual ||B - A X||_F / ||B||_F (:461-469)
        if (ctl_init(t, co, 0.0, 1)) return BKH_DEVICE_ERROR;
        HIP_OK(hipMemcpyAsync(tmp, B, sizeof(zd) * ms, hipMemcpyDeviceToDevice, st));
        launch_zspmm1(st, A, s, mone, X, one, tmp);
        hipLaunchKernelGGL(k_znorm2, dim3(Gv), dim3(256), 0, st, ms, (const zd *)tmp, part);
        launch_finalize<1>(st, part, Gv, co, ZFnbFin());
        LAUNCH_OK();
        if (ctl_fetch(t, co)) return BKH_DEVICE_ERROR;
        inf.true_res = t->ctl_host->fnb / hc.fnb;
        *info = inf;
    }
    return ret;
}

// Thin QR of an m x s complex panel on the device (CholeskyQR2): d_V <- Q, R to the
// host (s x s column-major, real positive diagonal). Replaces zqreco (myqr.c:41-76).
extern "C" int bkh_zqr(int m, int s, double *d_V, double *R) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!d_V || !R || m < 1 || s < 1 || s > ZB_SMAX || m < s) return BKH_BAD_ARGUMENT;
    const int Gg = zgram_grid(m);
    void *ws = bkh_workspace(carve_size(sizeof(Ctl)) + carve_size(sizeof(ZBState)) +
                             carve_size(sizeof(double) * Gg * ZB_SN * 2));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1);
    ZBState *bs = cv.take<ZBState>(1);
    double *part = cv.take<double>((size_t)Gg * ZB_SN * 2);
    hipStream_t st = t->stream;
    if (ctl_init(t, c, 0.0, 1)) return BKH_DEVICE_ERROR;
    zqr2(st, m, s, (zd *)d_V, bs, c, part, ZB_QR2_INIT);
    LAUNCH_OK();
    if (ctl_fetch(t, c)) return BKH_DEVICE_ERROR;
    if (t->ctl_host->status == ST_BREAKDOWN) {
        bkh_set_error("bkh_zqr: the panel is numerically rank deficient");
        return -2;
    }
    HIP_OK(hipMemcpy(R, bs->xi, sizeof(zd) * s * s, hipMemcpyDeviceToHost));
    return 0;
}
