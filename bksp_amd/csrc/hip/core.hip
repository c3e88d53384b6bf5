// core.hip -- runtime, memory, matrix upload and single-kernel entry points.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <new>
#include <stdlib.h>
#include "bkh_host.hpp"
#include "bkh_spmv.hpp"
#include "bkh_solver_common.hpp"

static thread_local char g_err[512];
static std::once_flag g_once;
static int g_init_rc = 0;
static thread_local ThreadCtx *g_tctx = nullptr;
static thread_local double g_last_ms = 0.0;

void bkh_set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

extern "C" const char *bkh_last_error(void) { return g_err; }

static void global_init() {
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev < 1) {
        bkh_set_error("no HIP device available (%s)", hipGetErrorString(e));
        g_init_rc = BKH_DEVICE_ERROR;
        return;
    }
    g_init_rc = 0;
}

extern "C" int bkh_init(void) {
    std::call_once(g_once, global_init);
    if (g_init_rc) bkh_set_error("no HIP device available");
    return g_init_rc;
}

std::atomic<int> g_bkh_fail_alloc{0};

extern "C" void bkh_debug_fail_alloc(int n) { g_bkh_fail_alloc.store(n > 0 ? n : 0); }
extern "C" int bkh_debug_fail_alloc_pending(void) { return std::max(0, g_bkh_fail_alloc.load()); }

ThreadCtx *bkh_tctx() {
    if (g_tctx) return g_tctx;
    if (bkh_init() != 0) return nullptr;
    ThreadCtx *t = new (std::nothrow) ThreadCtx();
    if (!t) return nullptr;
    if (hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void **)&t->ctl_host, sizeof(Ctl), hipHostMallocDefault) != hipSuccess ||
        hipEventCreate(&t->ev0) != hipSuccess || hipEventCreate(&t->ev1) != hipSuccess ||
        bkh_dmalloc((void **)&t->red2, sizeof(double) * RED2_CAP) != hipSuccess) {
        bkh_set_error("failed to create the thread's HIP stream/events");
        delete t;
        return nullptr;
    }
    g_tctx = t;
    return t;
}

double *bkh_red2() {
    ThreadCtx *t = bkh_tctx();
    return t ? t->red2 : nullptr;
}

__global__ __launch_bounds__(256) void k_part_reduce(const double *part, int nblocks, int E, double *out) {
    __shared__ double sh[256];
    const int t = threadIdx.x;
    const int b0 = blockIdx.x * RED_CHUNK, b1 = min(nblocks, b0 + RED_CHUNK);
    const int groups = E >= 256 ? 1 : 256 / E;
    for (int e0 = 0; e0 < E; e0 += 256) {
        const int e = e0 + (E >= 256 ? t : t % E), g = E >= 256 ? 0 : t / E;
        double acc = 0.0;
        if (e < E && g < groups)
            for (int b = b0 + g; b < b1; b += groups) acc += part[(size_t)b * E + e];
        sh[t] = acc;
        __syncthreads();
        if (t < 256 && (E >= 256 ? true : t < E) && e < E) {
            double v = 0.0;
            for (int q = 0; q < groups; q++) v += sh[(E >= 256 ? t : t % E) + q * (E >= 256 ? 0 : E)];
            out[(size_t)blockIdx.x * E + e] = v;
        }
        __syncthreads();
    }
}

void *bkh_workspace(size_t bytes) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return nullptr;
    if (bytes <= t->ws_bytes) return t->ws;
    if (t->ws) {
        hipStreamSynchronize(t->stream);
        bkh_graph_epoch_bump();
        hipFree(t->ws);
        t->ws = nullptr;
        t->ws_bytes = 0;
    }
    size_t want = bytes + bytes / 4;     // headroom: a regrowth invalidates cached graphs
    if (bkh_dmalloc(&t->ws, want) != hipSuccess) {
        bkh_set_error("workspace allocation of %zu bytes failed", want);
        t->ws = nullptr;
        return nullptr;
    }
    t->ws_bytes = want;
    return t->ws;
}

// ----------------------------------------------------------- graph cache

static std::atomic<uintptr_t> g_graph_epoch{1};
uintptr_t bkh_graph_epoch() { return g_graph_epoch.load(); }
void bkh_graph_epoch_bump() { g_graph_epoch.fetch_add(1); }

#define GRAPH_CACHE_MAX 24

hipGraphExec_t graph_cache_find(ThreadCtx *t, const GraphKey &key) {
    for (auto &e : t->graphs)
        if (e.first == key) return e.second;
    return nullptr;
}

void graph_cache_put(ThreadCtx *t, const GraphKey &key, hipGraphExec_t ge) {
    const uintptr_t ep = bkh_graph_epoch();
    // drop graphs of past epochs, then the oldest if still full (all work on the
    // thread's stream has completed before a new capture begins)
    std::vector<std::pair<std::vector<uintptr_t>, hipGraphExec_t>> keep;
    for (auto &e : t->graphs) {
        if (e.first.size() >= 4 && e.first[e.first.size() - 4] == ep) keep.push_back(e);
        else hipGraphExecDestroy(e.second);
    }
    if (keep.size() >= GRAPH_CACHE_MAX) {
        hipGraphExecDestroy(keep.front().second);
        keep.erase(keep.begin());
    }
    keep.emplace_back(key, ge);
    t->graphs.swap(keep);
}

extern "C" int bkh_device_name(char *buf, int len) {
    if (bkh_init() != 0) return BKH_DEVICE_ERROR;
    hipDeviceProp_t p;
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipGetDeviceProperties(&p, dev));
    snprintf(buf, (size_t)len, "%s %s CUs=%d", p.name, p.gcnArchName, p.multiProcessorCount);
    return 0;
}

extern "C" int bkh_sync(void) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    HIP_OK(hipStreamSynchronize(t->stream));
    return 0;
}

// ------------------------------------------------------------------ memory

extern "C" void *bkh_malloc(size_t bytes) {
    if (!bkh_tctx()) return nullptr;
    void *p = nullptr;
    if (bkh_dmalloc(&p, bytes ? bytes : 1) != hipSuccess) {
        bkh_set_error("bkh_dmalloc(%zu) failed", bytes);
        return nullptr;
    }
    return p;
}

extern "C" void bkh_free(void *p) {
    if (!p) return;
    ThreadCtx *t = bkh_tctx();
    if (t) hipStreamSynchronize(t->stream);
    hipFree(p);
}

extern "C" int bkh_h2d(void *d, const void *h, size_t bytes) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    HIP_OK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, t->stream));
    HIP_OK(hipStreamSynchronize(t->stream));
    return 0;
}

extern "C" int bkh_d2h(void *h, const void *d, size_t bytes) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    HIP_OK(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, t->stream));
    HIP_OK(hipStreamSynchronize(t->stream));
    return 0;
}

extern "C" int bkh_d2d(void *d, const void *s, size_t bytes) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    HIP_OK(hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, t->stream));
    HIP_OK(hipStreamSynchronize(t->stream));
    return 0;
}

__global__ void k_fill_bytes(unsigned char *p, unsigned char v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = v;
}

extern "C" int bkh_memset(void *d, int v, size_t bytes) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    // a fill kernel rather than hipMemsetAsync, so callers may capture it in graphs
    hipLaunchKernelGGL(k_fill_bytes, dim3(vec_grid((long long)bytes, 256, 16)), dim3(256), 0,
                       t->stream, (unsigned char *)d, (unsigned char)v, bytes);
    LAUNCH_OK();
    HIP_OK(hipStreamSynchronize(t->stream));
    return 0;
}

// ------------------------------------------------------------ control block

__global__ void k_ctl_init(Ctl *c, double tol, int max_it, int it0) {
    if (threadIdx.x == 0) {
        c->done = max_it <= it0 ? 1 : 0;
        c->status = max_it <= it0 ? ST_MAXIT : ST_RUNNING;
        c->it = it0; c->result_it = it0; c->max_it = max_it; c->inner_total = 0;
        c->base = it0; c->first = it0; c->tol = tol; c->fnb = 0; c->err = 0;
    }
    for (int i = threadIdx.x; i < BKH_MAX_SCAL; i += blockDim.x) c->s[i] = 0.0;
}

int ctl_init(ThreadCtx *t, Ctl *ctl_dev, double tol, int max_it, int it0) {
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(64), 0, t->stream, ctl_dev, tol, max_it, it0);
    LAUNCH_OK();
    return 0;
}

int ctl_fetch(ThreadCtx *t, const Ctl *ctl_dev) {
    HIP_OK(hipMemcpyAsync(t->ctl_host, ctl_dev, sizeof(Ctl), hipMemcpyDeviceToHost, t->stream));
    HIP_OK(hipStreamSynchronize(t->stream));
    return 0;
}

int ctl_result(const Ctl *c) {
    if (c->status == ST_CONVERGED) return c->result_it;
    if (c->status == ST_BREAKDOWN) return -2;   // DIVERGENCE
    return -1;                                    // NOT_CONVERGED
}

bool bkh_eager() {
    static const bool on = [] {
        const char *e = getenv("BKH_EAGER");
        return e && e[0] == '1';
    }();
    return on;
}

hipGraphExec_t eager_handle(ThreadCtx *t, std::function<void()> body, int K) {
    if (t->eager.size() >= 8) t->eager.erase(t->eager.begin());
    t->eager.emplace_back(new std::pair<std::function<void()>, int>(std::move(body), K));
    return (hipGraphExec_t)t->eager.back().get();
}

static std::pair<std::function<void()>, int> *eager_find(ThreadCtx *t, hipGraphExec_t ge) {
    for (auto &e : t->eager)
        if ((hipGraphExec_t)e.get() == ge) return e.get();
    return nullptr;
}

int run_graph_until_done(ThreadCtx *t, hipGraphExec_t ge, Ctl *ctl_dev, int max_launches) {
    if (bkh_eager()) {
        // in eager mode every handle is an iteration body, never a graph: a handle that
        // is no longer kept must not reach hipGraphLaunch
        auto *eb = eager_find(t, ge);
        if (!eb) {
            bkh_set_error("eager iteration body no longer held (more than 8 live bodies on this thread)");
            return BKH_DEVICE_ERROR;
        }
        for (int l = 1; l <= max_launches; l++) {
            for (int k = 0; k < eb->second; k++) eb->first();
            LAUNCH_OK();
            if (ctl_fetch(t, ctl_dev)) return BKH_DEVICE_ERROR;
            if (t->ctl_host->done) break;
        }
        return ctl_fetch(t, ctl_dev);
    }
    // One graph stays queued ahead of the check; kernels of a stopped solve return
    // at entry, so an extra replay costs only launch latency.
    HIP_OK(hipGraphLaunch(ge, t->stream));
    for (int l = 1; l <= max_launches; l++) {
        HIP_OK(hipMemcpyAsync(t->ctl_host, ctl_dev, sizeof(Ctl), hipMemcpyDeviceToHost, t->stream));
        hipEvent_t ev = t->ev1;
        HIP_OK(hipEventRecord(ev, t->stream));
        if (l < max_launches) HIP_OK(hipGraphLaunch(ge, t->stream));
        HIP_OK(hipEventSynchronize(ev));
        if (t->ctl_host->done) break;
    }
    return ctl_fetch(t, ctl_dev);        // its copy follows the queued replay on the stream
}

// --------------------------------------------------------------- matrices

static int check_csr(int n, const int *rp, const int *ci) {
    if (n < 1 || !rp || !ci) return BKH_BAD_ARGUMENT;
    if (rp[0] != 0) return BKH_BAD_ARGUMENT;
    for (int i = 0; i < n; i++)
        if (rp[i + 1] < rp[i]) return BKH_BAD_ARGUMENT;
    for (long long k = 0; k < rp[n]; k++)
        if (ci[k] < 0 || ci[k] >= n) return BKH_BAD_ARGUMENT;
    return 0;
}

// Builds the rows the device streams, in the order the reference accumulates:
//  mode 'n': rows as stored (src/matvec.c:15-28 gathers in stored order);
//  mode 'l': one stored triangle expanded; processing (j, k) in the order of
//            src/matvec.c:108-118, the gather term goes to row j and the mirrored
//            scatter term (col != j) to row col, each appended in processing order;
//  mode 't': A^T, entry (j, k) appended to row col in processing order
//            (src/matvec.c:38-45).
static void build_ordered(int n, const int *rp, const int *ci, const double *v, char mode,
                          std::vector<int> &orp, std::vector<int> &oci, std::vector<double> &ov) {
    long long nnz = rp[n];
    orp.assign((size_t)n + 1, 0);
    if (mode == 'n') {
        for (int i = 0; i <= n; i++) orp[i] = rp[i];
        oci.assign(ci, ci + nnz);
        ov.assign(v, v + nnz);
        return;
    }
    std::vector<long long> cnt((size_t)n + 1, 0);
    for (int j = 0; j < n; j++)
        for (int k = rp[j]; k < rp[j + 1]; k++) {
            if (mode == 'l') {
                cnt[j]++;
                if (ci[k] != j) cnt[ci[k]]++;
            } else {
                cnt[ci[k]]++;
            }
        }
    long long tot = 0;
    for (int i = 0; i < n; i++) { orp[i] = (int)tot; tot += cnt[i]; }
    orp[n] = (int)tot;
    oci.resize((size_t)tot);
    ov.resize((size_t)tot);
    std::vector<long long> cur(orp.begin(), orp.end() - 1);
    for (int j = 0; j < n; j++)
        for (int k = rp[j]; k < rp[j + 1]; k++) {
            if (mode == 'l') {
                if (ci[k] != j) { long long p = cur[ci[k]]++; oci[p] = j; ov[p] = v[k]; }
                long long q = cur[j]++; oci[q] = ci[k]; ov[q] = v[k];
            } else {
                long long p = cur[ci[k]]++; oci[p] = j; ov[p] = v[k];
            }
        }
}

// Greedy row blocks of <= max_rows rows and <= cap nonzeros; a row longer than
// cap forms a block of its own.
static void build_blocks(int n, const std::vector<int> &rp, int max_rows, int cap,
                         std::vector<int> &blk) {
    blk.clear();
    blk.push_back(0);
    int r = 0;
    while (r < n) {
        int len = rp[r + 1] - rp[r];
        if (len > cap) { r++; blk.push_back(r); continue; }
        int start = r, nz = 0;
        while (r < n && r - start < max_rows) {
            int l = rp[r + 1] - rp[r];
            if (l > cap || nz + l > cap) break;
            nz += l;
            r++;
        }
        blk.push_back(r);
    }
}

static void free_matrix(bkh_matrix *A) {
    if (!A) return;
    hipFree(A->rp); hipFree(A->ci); hipFree(A->val); hipFree(A->val32);
    hipFree(A->diag); hipFree(A->blk); hipFree(A->wblk); hipFree(A->cd16);
    hipFree(A->sso); hipFree(A->sval); hipFree(A->sval32); hipFree(A->scol); hipFree(A->scd);
    hipFree(A->dso); hipFree(A->dval); hipFree(A->dval32); hipFree(A->doff); hipFree(A->dmeta);
    if (A->T) free_matrix(A->T);
    delete A;
}

// SELL-64 copy of the matrix (see bkh_host.hpp) when its padding is small; the
// default SpMV kernel for such matrices. Returns false only on a device error.
static bool upload_sell(bkh_matrix *A, int n, const std::vector<int> &rp, const std::vector<int> &ci,
                        const std::vector<double> &v, int want_fp32, hipStream_t s) {
    const int nsl = (n + 63) / 64;
    std::vector<int> so((size_t)nsl + 1, 0);
    long long tot = 0;
    for (int q = 0; q < nsl; q++) {
        int w = 0;
        for (int r = q * 64; r < std::min(n, q * 64 + 64); r++) w = std::max(w, rp[r + 1] - rp[r]);
        tot += 64LL * w;
        if (tot >= (1LL << 31) - 64) return true;             // keep CSR only
        so[q + 1] = (int)tot;
    }
    if ((double)tot > SELL_MAX_PAD * (double)std::max(1LL, A->nnz) || A->nnz == 0) return true;
    std::vector<double> sv((size_t)tot + 4, 0.0);
    std::vector<int> sc((size_t)tot + 4, 0);
    std::vector<short> sd((size_t)tot + 4, 0);
    bool fits = true;
    for (int q = 0; q < nsl; q++)
        for (int r = q * 64; r < std::min(n, q * 64 + 64); r++) {
            const int l = r & 63, w = (so[q + 1] - so[q]) / 64;
            for (int k = 0; k < w; k++) {
                const size_t e = (size_t)so[q] + 64 * (size_t)k + l;
                if (k < rp[r + 1] - rp[r]) {
                    const int c = ci[rp[r] + k];
                    sv[e] = v[rp[r] + k];
                    sc[e] = c;
                    const int d = c - r;
                    if (d < -32768 || d > 32767) fits = false;
                    sd[e] = (short)d;
                } else {
                    sc[e] = r;                                   // padding: a valid index, never read
                }
            }
        }
    const size_t ne = (size_t)tot + 4;
    bool ok = bkh_dmalloc(&A->sso, sizeof(int) * (nsl + 1)) == hipSuccess &&
              bkh_dmalloc(&A->sval, sizeof(double) * ne) == hipSuccess &&
              bkh_dmalloc(&A->scol, sizeof(int) * ne) == hipSuccess;
    if (ok && fits) ok = bkh_dmalloc(&A->scd, sizeof(short) * ne) == hipSuccess;
    if (ok && want_fp32) ok = bkh_dmalloc(&A->sval32, sizeof(float) * ne) == hipSuccess;
    if (!ok) return false;
    ok = hipMemcpyAsync(A->sso, so.data(), sizeof(int) * (nsl + 1), hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(A->sval, sv.data(), sizeof(double) * ne, hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(A->scol, sc.data(), sizeof(int) * ne, hipMemcpyHostToDevice, s) == hipSuccess;
    if (ok && fits) ok = hipMemcpyAsync(A->scd, sd.data(), sizeof(short) * ne, hipMemcpyHostToDevice, s) == hipSuccess;
    if (ok && want_fp32) {
        std::vector<float> s32(ne);
        for (size_t e = 0; e < ne; e++) s32[e] = (float)sv[e];
        ok = hipMemcpyAsync(A->sval32, s32.data(), sizeof(float) * ne, hipMemcpyHostToDevice, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
    }
    ok = ok && hipStreamSynchronize(s) == hipSuccess;
    if (!ok) return false;
    A->nsl = nsl;
    A->sell_entries = tot;
    A->variant = SPMV_SELL;
    return true;
}

// Slice-diagonal copy of a SELL matrix (see bkh_host.hpp) when it qualifies;
// BKH_DIA=0 in the environment turns it off. Returns false only on a device error.
static bool upload_dia(bkh_matrix *A, int n, const std::vector<int> &rp, const std::vector<int> &ci,
                       const std::vector<double> &v, int want_fp32, hipStream_t s) {
    const char *env = getenv("BKH_DIA");
    if (env && env[0] == '0') return true;
    const int nsl = A->nsl;
    if (nsl <= 0 || A->nnz == 0) return true;
    for (int r = 0; r < n; r++)
        for (int k = rp[r] + 1; k < rp[r + 1]; k++)
            if (ci[k] <= ci[k - 1]) return true;                   // a row not in ascending order
    std::vector<int> so((size_t)nsl + 1, 0), offs;
    std::vector<int> u;
    long long tot = 0;
    const long long cap = (long long)(DIA_MAX_PAD * (double)A->nnz) + 64;
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
    std::vector<double> dv((size_t)tot + 4, 0.0);
    for (int q = 0; q < nsl; q++) {
        const int *o = offs.data() + so[q] / 64;
        for (int r = q * 64; r < std::min(n, q * 64 + 64); r++) {
            int p = 0;
            for (int k = rp[r]; k < rp[r + 1]; k++) {
                while (o[p] != ci[k] - r) p++;                     // both ascend; o holds every offset
                dv[(size_t)so[q] + 64 * (size_t)p + (r & 63)] = v[k];
                p++;
            }
        }
    }
    const size_t ne = (size_t)tot + 4;
    bool ok = bkh_dmalloc(&A->dso, sizeof(int) * (nsl + 1)) == hipSuccess &&
              bkh_dmalloc(&A->dval, sizeof(double) * ne) == hipSuccess &&
              bkh_dmalloc(&A->doff, sizeof(int) * (offs.size() + 8)) == hipSuccess;
    if (ok && want_fp32) ok = bkh_dmalloc(&A->dval32, sizeof(float) * ne) == hipSuccess;
    if (!ok) return false;
    offs.resize(offs.size() + 8, 0);                              // read-ahead padding of the kernels
    ok = hipMemcpyAsync(A->dso, so.data(), sizeof(int) * (nsl + 1), hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(A->dval, dv.data(), sizeof(double) * ne, hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(A->doff, offs.data(), sizeof(int) * offs.size(), hipMemcpyHostToDevice, s) == hipSuccess;
    std::vector<float> d32;
    if (ok && want_fp32) {
        d32.resize(ne);
        for (size_t e = 0; e < ne; e++) d32[e] = (float)dv[e];
        ok = hipMemcpyAsync(A->dval32, d32.data(), sizeof(float) * ne, hipMemcpyHostToDevice, s) == hipSuccess;
    }
    // per-slice records for the pipelined block kernels
    std::vector<DiaSlice> meta((size_t)nsl);
    bool qualifies = true;
    int maxpos = 0;
    for (int q = 0; q < nsl && qualifies; q++) {
        DiaSlice &m = meta[q];
        m.s0 = so[q];
        m.len = (so[q + 1] - so[q]) / 64;
        maxpos = std::max(maxpos, m.len);
        if (m.len < 1 || m.len > DIA_META_POS) { qualifies = false; break; }
        const int *o = offs.data() + so[q] / 64;
        m.cq = -1;
        for (int k = 0; k < DIA_META_POS; k++) m.o[k] = o[std::min(k, m.len - 1)];
        for (int k = 0; k < m.len; k++)
            if (o[k] == 0) m.cq = k;
        if (m.cq < 0) qualifies = false;
        const long long lo = 64LL * q + o[0], hi = 64LL * q + 63 + o[m.len - 1];
        m.inter = (lo >= 0 && hi <= (long long)n - 1) ? 1 : 0;
    }
    if (ok && qualifies) {
        ok = bkh_dmalloc(&A->dmeta, sizeof(DiaSlice) * nsl) == hipSuccess &&
             hipMemcpyAsync(A->dmeta, meta.data(), sizeof(DiaSlice) * nsl, hipMemcpyHostToDevice, s) == hipSuccess;
    }
    ok = ok && hipStreamSynchronize(s) == hipSuccess;
    if (!ok) return false;
    A->dia_entries = tot;
    A->dia_maxpos = maxpos;
    int maxoff = 0;
    for (size_t e = 0; e + 8 < offs.size(); e++) maxoff = std::max(maxoff, std::abs(offs[e]));
    A->dia_maxoff = maxoff;
    return true;
}

static bkh_matrix *upload(int n, const std::vector<int> &rp, const std::vector<int> &ci,
                          const std::vector<double> &v, int want_fp32) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return nullptr;
    bkh_matrix *A = new (std::nothrow) bkh_matrix();
    if (!A) return nullptr;
    A->n = n;
    A->nnz = rp[n];
    std::vector<int> blk, wblk;
    build_blocks(n, rp, SPMV_ROWS, SPMV_CAP, blk);
    build_blocks(n, rp, SPMV_WROWS, SPMV_WCAP, wblk);
    A->nblk = (int)blk.size() - 1;
    A->nwblk = (int)wblk.size() - 1;
    for (int i = 0; i < n; i++) A->max_row = std::max(A->max_row, rp[i + 1] - rp[i]);
    // the wave variant serves a row longer than SPMV_WCAP with a single wave; keep
    // the workgroup variant for matrices with such rows
    A->variant = (A->max_row <= SPMV_WCAP) ? SPMV_WAVE : SPMV_BLOCK;
    std::vector<double> d((size_t)n, 0.0);
    for (int i = 0; i < n; i++)
        for (int k = rp[i]; k < rp[i + 1]; k++)
            if (ci[k] == i) d[i] += v[k];
    // +4 elements: the SpMV's 16-byte vector loads start at an aligned index and
    // may cover up to 3 entries past the last nonzero (masked, never used).
    size_t nz = (size_t)(A->nnz > 0 ? A->nnz : 1) + 4;
    bool ok = bkh_dmalloc(&A->rp, sizeof(int) * (n + 1)) == hipSuccess &&
              bkh_dmalloc(&A->ci, sizeof(int) * nz) == hipSuccess &&
              bkh_dmalloc(&A->val, sizeof(double) * nz) == hipSuccess &&
              bkh_dmalloc(&A->diag, sizeof(double) * n) == hipSuccess &&
              bkh_dmalloc(&A->blk, sizeof(int) * blk.size()) == hipSuccess &&
              bkh_dmalloc(&A->wblk, sizeof(int) * wblk.size()) == hipSuccess;
    if (ok && want_fp32) ok = bkh_dmalloc(&A->val32, sizeof(float) * nz) == hipSuccess;
    if (!ok) {
        bkh_set_error("device allocation for a %d-row matrix failed", n);
        free_matrix(A);
        return nullptr;
    }
    hipStream_t s = t->stream;
    ok = hipMemcpyAsync(A->rp, rp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(A->ci, ci.data(), sizeof(int) * A->nnz, hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(A->val, v.data(), sizeof(double) * A->nnz, hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(A->diag, d.data(), sizeof(double) * n, hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(A->blk, blk.data(), sizeof(int) * blk.size(), hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(A->wblk, wblk.data(), sizeof(int) * wblk.size(), hipMemcpyHostToDevice, s) == hipSuccess;
    std::vector<float> v32;
    if (ok && want_fp32) {
        // 16-bit column deltas relative to the first row of the entry's wave block
        std::vector<short> cd((size_t)A->nnz);
        bool fits = true;
        for (size_t b = 0; b + 1 < wblk.size() && fits; b++) {
            const int r0 = wblk[b];
            for (int r = wblk[b]; r < wblk[b + 1] && fits; r++)
                for (int k = rp[r]; k < rp[r + 1]; k++) {
                    const int d = ci[k] - r0;
                    if (d < -32768 || d > 32767) { fits = false; break; }
                    cd[k] = (short)d;
                }
        }
        if (fits && A->max_row <= SPMV_WCAP) {
            ok = bkh_dmalloc(&A->cd16, sizeof(short) * nz) == hipSuccess &&
                 hipMemcpyAsync(A->cd16, cd.data(), sizeof(short) * A->nnz, hipMemcpyHostToDevice, s) == hipSuccess &&
                 hipStreamSynchronize(s) == hipSuccess;
        }
    }
    if (ok && want_fp32) {
        v32.resize((size_t)A->nnz);
        for (long long k = 0; k < A->nnz; k++) v32[k] = (float)v[k];
        ok = hipMemcpyAsync(A->val32, v32.data(), sizeof(float) * A->nnz, hipMemcpyHostToDevice, s) == hipSuccess;
    }
    if (ok) ok = upload_sell(A, n, rp, ci, v, want_fp32, s);
    if (ok) ok = upload_dia(A, n, rp, ci, v, want_fp32, s);
    if (!ok || hipStreamSynchronize(s) != hipSuccess) {
        bkh_set_error("matrix upload failed");
        free_matrix(A);
        return nullptr;
    }
    return A;
}

extern "C" bkh_matrix *bkh_matrix_create(int n, const int *row_ptr, const int *col,
                                         const double *val, int stored_triangle,
                                         int want_transpose, int want_fp32) {
    if (check_csr(n, row_ptr, col) != 0 || (!val && row_ptr[n] > 0)) {
        bkh_set_error("invalid CSR input (n=%d)", n);
        return nullptr;
    }
    std::vector<int> orp, oci;
    std::vector<double> ov;
    build_ordered(n, row_ptr, col, val, stored_triangle ? 'l' : 'n', orp, oci, ov);
    bkh_matrix *A = upload(n, orp, oci, ov, want_fp32);
    if (!A) return nullptr;
    A->stored_triangle = stored_triangle;
    if (want_transpose) {
        if (stored_triangle) {
            A->T = upload(n, orp, oci, ov, want_fp32);   // symmetric: A^T = A
        } else {
            std::vector<int> trp, tci;
            std::vector<double> tv;
            build_ordered(n, row_ptr, col, val, 't', trp, tci, tv);
            A->T = upload(n, trp, tci, tv, want_fp32);
        }
        if (!A->T) { free_matrix(A); return nullptr; }
    }
    return A;
}

extern "C" void bkh_matrix_destroy(bkh_matrix *A) {
    ThreadCtx *t = bkh_tctx();
    if (t) hipStreamSynchronize(t->stream);
    bkh_graph_epoch_bump();
    free_matrix(A);
}

extern "C" long long bkh_matrix_nnz(const bkh_matrix *A) { return A ? A->nnz : -1; }
extern "C" long long bkh_matrix_dia_entries(const bkh_matrix *A) { return A ? A->dia_entries : -1; }
extern "C" int bkh_matrix_rows(const bkh_matrix *A) { return A ? A->n : -1; }
extern "C" int bkh_matrix_blocks(const bkh_matrix *A) { return A ? spmv_grid(A) : -1; }

// 0: workgroup CSR-stream, 1: wave CSR-stream, 2: SELL-64 (16-bit column deltas
// when they fit), 3: SELL-64 with 32-bit columns. SELL needs the matrix to have a
// SELL copy (small padding); the wave variant needs rows of <= SPMV_WCAP nonzeros.
static int set_variant(bkh_matrix *A, int variant) {
    if (variant == SPMV_WAVE && A->max_row > SPMV_WCAP) return BKH_BAD_ARGUMENT;
    if ((variant == SPMV_SELL || variant == SPMV_SELL32) && A->nsl == 0) return BKH_BAD_ARGUMENT;
    if (variant < SPMV_BLOCK || variant > SPMV_SELL32) return BKH_BAD_ARGUMENT;
    A->variant = variant;
    bkh_graph_epoch_bump();
    return 0;
}

extern "C" int bkh_matrix_set_spmv(bkh_matrix *A, int variant) {
    if (!A) return BKH_BAD_ARGUMENT;
    int rc = set_variant(A, variant);
    if (rc) { bkh_set_error("SpMV variant %d is not available for this matrix", variant); return rc; }
    if (A->T) set_variant(A->T, variant);
    return 0;
}

// ------------------------------------------------------------- autotuner

// The wave variant's blocks for `rows` rows / `cap` nonzeros per wavefront, with the
// 16-bit column deltas (relative to a block's first row) rebuilt for them when the
// matrix keeps them (dropped if a delta no longer fits). 0 or a device error.
static int set_wave_blocking(bkh_matrix *A, const std::vector<int> &rp, const std::vector<int> &ci, int rows,
                             int cap) {
    if (rows == A->wrows && cap == A->wcap) return 0;
    std::vector<int> wb;
    build_blocks(A->n, rp, rows, cap, wb);
    int *d = nullptr;
    if (bkh_dmalloc(&d, sizeof(int) * wb.size()) != hipSuccess ||
        hipMemcpy(d, wb.data(), sizeof(int) * wb.size(), hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(d);
        bkh_set_error("autotuner: wave blocks upload failed");
        return BKH_DEVICE_ERROR;
    }
    bkh_graph_epoch_bump();                       // cached graphs hold the old blocks
    hipFree(A->wblk);
    A->wblk = d;
    A->nwblk = (int)wb.size() - 1;
    A->wrows = rows;
    A->wcap = cap;
    if (A->cd16) {
        std::vector<short> cd((size_t)A->nnz);
        bool fits = true;
        for (size_t b = 0; b + 1 < wb.size() && fits; b++)
            for (int r = wb[b]; r < wb[b + 1] && fits; r++)
                for (int k = rp[r]; k < rp[r + 1]; k++) {
                    const int dd = ci[k] - wb[b];
                    if (dd < -32768 || dd > 32767) { fits = false; break; }
                    cd[k] = (short)dd;
                }
        if (!fits) {
            hipFree(A->cd16);
            A->cd16 = nullptr;
        } else if (hipMemcpy(A->cd16, cd.data(), sizeof(short) * A->nnz, hipMemcpyHostToDevice) != hipSuccess) {
            bkh_set_error("autotuner: column deltas upload failed");
            return BKH_DEVICE_ERROR;
        }
    }
    return 0;
}

extern "C" int bkh_matrix_tune(bkh_matrix *A, int max_s, bkh_matrix_tuning *out) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !out || max_s < 0 || max_s > 16) return BKH_BAD_ARGUMENT;
    memset(out, 0, sizeof *out);
    hipStream_t s = t->stream;
    HIP_OK(hipStreamSynchronize(s));
    const int n = A->n;
    std::vector<int> rp((size_t)n + 1), ci;
    HIP_OK(hipMemcpy(rp.data(), A->rp, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
    if (A->cd16) {
        ci.resize((size_t)A->nnz);
        HIP_OK(hipMemcpy(ci.data(), A->ci, sizeof(int) * A->nnz, hipMemcpyDeviceToHost));
    }
    const int ms = std::max(1, max_s);
    void *ws = bkh_workspace(4 * carve_size(sizeof(double) * (size_t)n * ms));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    double *x = cv.take<double>((size_t)n * ms), *y = cv.take<double>((size_t)n * ms);
    HIP_OK(hipMemsetAsync(x, 0, sizeof(double) * n * ms, s));
    HIP_OK(hipMemsetAsync(y, 0, sizeof(double) * n * ms, s));
    const int reps = 5;
    auto timed = [&](auto &&f) -> double {
        for (int r = 0; r < reps + 1; r++) {
            if (r == 1) hipEventRecord(t->ev0, s);
            f();
        }
        hipEventRecord(t->ev1, s);
        hipEventSynchronize(t->ev1);
        float msv = 0.f;
        hipEventElapsedTime(&msv, t->ev0, t->ev1);
        return msv / reps;
    };
    Epi<0> e0;
    memset(&e0, 0, sizeof e0);
    auto spmv_once = [&]() {
        launch_spmv<double, double, 0>(s, A, A->val, x, y, 1.0, 0.0, e0, nullptr, nullptr, NoFin(), true);
    };
    // SpMV candidates {variant, rows, cap}
    struct Cand { int v, rows, cap; };
    std::vector<Cand> cand;
    if (A->nsl > 0) {
        cand.push_back({SPMV_SELL, A->wrows, A->wcap});
        if (A->scd) cand.push_back({SPMV_SELL32, A->wrows, A->wcap});
    }
    if (A->max_row <= SPMV_WCAP && n > 0) {
        std::vector<int> len(n);
        long long tot = 0;
        for (int i = 0; i < n; i++) { len[i] = rp[i + 1] - rp[i]; tot += len[i]; }
        std::sort(len.begin(), len.end());
        const double p50 = len[n / 2], p90 = len[(size_t)n * 9 / 10], mean = (double)tot / n;
        for (int rows : {16, 32, 64})
            for (double per : {p50, p90, mean, (double)SPMV_WCAP / rows}) {
                int cap = (int)std::ceil(rows * per / 64.0) * 64;
                cap = std::min(SPMV_WCAP, std::max(cap, std::max(A->max_row, 64)));
                bool dup = false;
                for (const Cand &c : cand) dup |= (c.v == SPMV_WAVE && c.rows == rows && c.cap == cap);
                if (!dup) cand.push_back({SPMV_WAVE, rows, cap});
            }
    }
    cand.push_back({SPMV_BLOCK, A->wrows, A->wcap});
    const int v0 = A->variant, r0 = A->wrows, c0 = A->wcap;
    out->spmv_ms_default = timed(spmv_once);
    double best = 1e300;
    Cand bc = {v0, r0, c0};
    for (const Cand &c : cand) {
        if (c.v == SPMV_WAVE) {
            const int rc = set_wave_blocking(A, rp, ci.empty() ? rp : ci, c.rows, c.cap);
            if (rc) return rc;
        }
        A->variant = c.v;
        const double msv = timed(spmv_once);
        if (msv < best) { best = msv; bc = c; }
    }
    A->variant = bc.v;
    {
        const int rc = set_wave_blocking(A, rp, ci.empty() ? rp : ci, bc.v == SPMV_WAVE ? bc.rows : r0,
                                         bc.v == SPMV_WAVE ? bc.cap : c0);
        if (rc) return rc;
    }
    bkh_graph_epoch_bump();
    out->spmv_variant = A->variant;
    out->rows_per_wave = A->wrows;
    out->nnz_per_wave = A->wcap;
    out->candidates = (int)cand.size();
    out->spmv_ms = best;
    // bkh_spmm's columns per pass for each s
    for (int sc = 1; sc <= max_s; sc++) {
        double bw = 1e300;
        int wbest = sc;
        for (int w : {1, 2, 4, 8, 16, sc}) {
            if (w > sc) continue;
            int rc = 0;
            const double msv = timed([&]() { rc |= spmm_chunked(t, A, sc, w, 1.0, x, 0.0, y); });
            if (rc) return rc;
            if (w == sc) out->spmm_ms_full[sc] = msv;
            if (msv < bw) { bw = msv; wbest = w; }
        }
        A->spmm_w[sc] = wbest;
        out->spmm_width[sc] = wbest;
        out->spmm_ms[sc] = bw;
    }
    LAUNCH_OK();
    HIP_OK(hipStreamSynchronize(s));
    return 0;
}

// ------------------------------------------------------------- single kernels

extern "C" int bkh_spmv(const bkh_matrix *A, char op, double alpha, const double *x, double beta,
                        double *y) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !x || !y) return BKH_BAD_ARGUMENT;
    const bkh_matrix *M = A;
    if (op == 't') {
        if (!A->T) { bkh_set_error("matrix was created without its transpose"); return BKH_BAD_ARGUMENT; }
        M = A->T;
    } else if (op != 'n') {
        return BKH_BAD_ARGUMENT;
    }
    Epi<0> e;
    memset(&e, 0, sizeof e);
    launch_spmv<double, double, 0>(t->stream, M, M->val, x, y, alpha, beta, e, nullptr, nullptr, NoFin(), false);
    LAUNCH_OK();
    HIP_OK(hipStreamSynchronize(t->stream));
    return 0;
}

extern "C" int bkh_spmv32(const bkh_matrix *A, const float *x, float *y) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !A->val32 || !x || !y) {
        bkh_set_error("bkh_spmv32 needs a matrix created with want_fp32");
        return BKH_BAD_ARGUMENT;
    }
    Epi<0, float> e;
    memset(&e, 0, sizeof e);
    launch_spmv<float, float, 0>(t->stream, A, A->val32, x, y, 1.0, 0.0, e, nullptr, nullptr, NoFin(), false);
    LAUNCH_OK();
    HIP_OK(hipStreamSynchronize(t->stream));
    return 0;
}

// Deterministic dot: one partial per workgroup, fixed-order final sum.
struct StoreFin {
    __device__ void operator()(Ctl *c, const double *tot) const { c->s[0] = tot[0]; }
};

__global__ __launch_bounds__(256) void k_dot(long long n, const double *x, const double *y, Ctl *c,
                                             double *part) {
    double v[1] = {0.0};
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        v[0] += x[i] * y[i];
    block_partial<1, 256>(v, part);
}

extern "C" int bkh_dot(long long n, const double *x, const double *y, double *out) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    int G = vec_grid(n, 256);
    void *ws = bkh_workspace(carve_size(sizeof(Ctl)) + carve_size(sizeof(double) * G));
    if (!ws) return BKH_DEVICE_ERROR;
    Carve cv(ws);
    Ctl *c = cv.take<Ctl>(1);
    double *part = cv.take<double>(G);
    if (ctl_init(t, c, 0.0, 1)) return BKH_DEVICE_ERROR;
    hipLaunchKernelGGL(k_dot, dim3(G), dim3(256), 0, t->stream, n, x, y, c, part);
    launch_finalize<1>(t->stream, part, G, c, StoreFin());
    LAUNCH_OK();
    if (ctl_fetch(t, c)) return BKH_DEVICE_ERROR;
    *out = t->ctl_host->s[0];
    return 0;
}

extern "C" double bkh_last_solve_ms(void) { return g_last_ms; }
void bkh_note_solve_ms(double ms) { g_last_ms = ms; }

// ------------------------------------------------- residuals and refinement

struct ResidFin {
    __device__ void operator()(Ctl *c, const double *tot) const {
        double rn = sqrt(tot[0]), bn = sqrt(tot[1]);
        c->s[0] = rn / bn;
        c->s[1] = rn;
        c->s[2] = bn;
    }
};

// r = b - y - sigma x, where r holds y = A x on entry.
__global__ __launch_bounds__(256) void k_resid(long long n, const double *b, const double *x,
                                               double sigma, double *r, Ctl *c, double *part) {
    double v[2] = {0.0, 0.0};
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        double ri = b[i] - r[i];
        if (sigma != 0.0) ri -= sigma * x[i];
        r[i] = ri;
        v[0] += ri * ri;
        v[1] += b[i] * b[i];
    }
    block_partial<2, 256>(v, part);
}

int resid_fp64(ThreadCtx *t, const bkh_matrix *A, double sigma, const double *b, const double *x,
               double *r, Ctl *c, double *part) {
    Epi<0> e;
    memset(&e, 0, sizeof e);
    launch_spmv<double, double, 0>(t->stream, A, A->val, x, r, 1.0, 0.0, e, nullptr, nullptr, NoFin());
    const int G = vec_grid(A->n, 256);
    hipLaunchKernelGGL(k_resid, dim3(G), dim3(256), 0, t->stream, (long long)A->n, b, x, sigma, r, c, part);
    launch_finalize<2>(t->stream, part, G, c, ResidFin());
    LAUNCH_OK();
    return 0;
}

int true_residual(ThreadCtx *t, const bkh_matrix *A, double sigma, const double *b, const double *x,
                  double *out) {
    const int G = vec_grid(A->n, 256);
    void *mem = nullptr;
    size_t bytes = carve_size(sizeof(Ctl)) + carve_size(sizeof(double) * G * 2) +
                   carve_size(sizeof(double) * A->n);
    HIP_OK(bkh_dmalloc(&mem, bytes));
    Carve cv(mem);
    Ctl *c = cv.take<Ctl>(1);
    double *part = cv.take<double>((size_t)G * 2);
    double *r = cv.take<double>(A->n);
    int rc = ctl_init(t, c, 0.0, 1);
    if (!rc) rc = resid_fp64(t, A, sigma, b, x, r, c, part);
    if (!rc) rc = ctl_fetch(t, c);
    if (!rc) *out = t->ctl_host->s[0];
    hipFree(mem);
    return rc;
}

__global__ __launch_bounds__(256) void k_ir_prep(long long n, const double *r64, const Ctl *c,
                                                 float *r32, float *d32) {
    const double inv = 1.0 / c->s[1];
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        r32[i] = (float)(r64[i] * inv);
        d32[i] = 0.0f;
    }
}

__global__ __launch_bounds__(256) void k_ir_update(long long n, double *x, const float *d32,
                                                   const Ctl *c) {
    const double sc = c->s[1];
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        x[i] += sc * (double)d32[i];
}

int ir_prep(ThreadCtx *t, long long n, const double *r64, const Ctl *c, float *r32, float *d32) {
    hipLaunchKernelGGL(k_ir_prep, dim3(vec_grid(n, 256)), dim3(256), 0, t->stream, n, r64, c, r32, d32);
    LAUNCH_OK();
    return 0;
}

int ir_update(ThreadCtx *t, long long n, double *x, const float *d32, const Ctl *c) {
    hipLaunchKernelGGL(k_ir_update, dim3(vec_grid(n, 256)), dim3(256), 0, t->stream, n, x, d32, c);
    LAUNCH_OK();
    return 0;
}

// Times `reps` back-to-back SpMV launches on this thread's stream with HIP events
// (the stream the kernels run on), returning the average milliseconds per launch.
// Bytes one SpMV with the matrix's current kernel must move: stored entries x
// (value + index bytes), row pointers, x read once and y written once.
extern "C" long long bkh_matrix_spmv_bytes(const bkh_matrix *A, int fp32) {
    if (!A) return -1;
    const long long vb = fp32 ? 4 : 8, n = A->n;
    if ((A->variant == SPMV_SELL || A->variant == SPMV_SELL32) && A->nsl > 0) {
        const long long ib = (A->variant == SPMV_SELL && A->scd) ? 2 : 4;
        return A->sell_entries * (vb + ib) + 4 * (n + 1) + 2 * vb * n;
    }
    return A->nnz * (vb + 4) + 4 * (n + 1) + 2 * vb * n;
}

extern "C" int bkh_bench_spmv(const bkh_matrix *A, int fp32, const void *x, void *y, int reps,
                              double *avg_ms) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (!A || !x || !y || reps < 1 || (fp32 && !A->val32)) return BKH_BAD_ARGUMENT;
    HIP_OK(hipEventRecord(t->ev0, t->stream));
    for (int r = 0; r < reps; r++) {
        if (fp32) {
            Epi<0, float> e;
            memset(&e, 0, sizeof e);
            launch_spmv<float, float, 0>(t->stream, A, A->val32, (const float *)x, (float *)y, 1.0, 0.0,
                                         e, nullptr, nullptr, NoFin(), false);
        } else {
            Epi<0> e;
            memset(&e, 0, sizeof e);
            launch_spmv<double, double, 0>(t->stream, A, A->val, (const double *)x, (double *)y, 1.0,
                                           0.0, e, nullptr, nullptr, NoFin(), false);
        }
    }
    LAUNCH_OK();
    HIP_OK(hipEventRecord(t->ev1, t->stream));
    HIP_OK(hipEventSynchronize(t->ev1));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    *avg_ms = ms / reps;
    return 0;
}

__global__ void k_copy(const double4 *__restrict__ a, double4 *__restrict__ b, long long n4) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
         i += (long long)gridDim.x * blockDim.x)
        b[i] = a[i];
}

// Device-copy bandwidth probe: b = a over `bytes` (multiple of 32), averaged.
extern "C" int bkh_bench_copy(const void *a, void *b, size_t bytes, int reps, double *avg_ms) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    long long n4 = (long long)(bytes / 32);
    HIP_OK(hipEventRecord(t->ev0, t->stream));
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, t->stream, (const double4 *)a, (double4 *)b, n4);
    LAUNCH_OK();
    HIP_OK(hipEventRecord(t->ev1, t->stream));
    HIP_OK(hipEventSynchronize(t->ev1));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    *avg_ms = ms / reps;
    return 0;
}

int finish_solve(ThreadCtx *t, const bkh_matrix *A, double sigma, const double *b, const double *x,
                 bkh_info *inf, bkh_info *info) {
    HIP_OK(hipEventRecord(t->ev1, t->stream));
    HIP_OK(hipEventSynchronize(t->ev1));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    inf->time_ms = ms;
    bkh_note_solve_ms(ms);
    const Ctl *c = t->ctl_host;
    inf->ret = ctl_result(c);
    inf->status = c->status;
    inf->iters = c->result_it;
    inf->rec_res = c->err;
    if (info) {
        double tr = 0.0;
        if (true_residual(t, A, sigma, b, x, &tr) == 0) inf->true_res = tr;
        *info = *inf;
    }
    return inf->ret;
}

// ||A||_F with a stored triangle's off-diagonals counted twice (src/norm.c:5-47).
__global__ __launch_bounds__(256) void k_spnorm(int n, const int *rp, const int *ci, const double *v,
                                                int sym, double *part) {
    double acc[1] = {0.0};
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        for (int k = rp[i]; k < rp[i + 1]; k++) {
            double t = v[k] * v[k];
            acc[0] += (sym && ci[k] != i) ? 2.0 * t : t;
        }
    block_partial<1, 256>(acc, part);
}

extern "C" int bkh_sp_norm_f(int n, const int *row_ptr, const int *col, const double *val,
                             int symmetric, double *out) {
    ThreadCtx *t = bkh_tctx();
    if (!t) return BKH_DEVICE_ERROR;
    if (check_csr(n, row_ptr, col) != 0) return BKH_BAD_ARGUMENT;
    const long long nnz = row_ptr[n];
    const int G = vec_grid(n, 256);
    void *mem = nullptr;
    size_t bytes = carve_size(sizeof(Ctl)) + carve_size(sizeof(double) * G) + carve_size(sizeof(int) * (n + 1)) +
                   carve_size(sizeof(int) * (nnz + 1)) + carve_size(sizeof(double) * (nnz + 1));
    HIP_OK(bkh_dmalloc(&mem, bytes));
    Carve cv(mem);
    Ctl *c = cv.take<Ctl>(1);
    double *part = cv.take<double>(G);
    int *drp = cv.take<int>(n + 1), *dci = cv.take<int>(nnz + 1);
    double *dv = cv.take<double>(nnz + 1);
    int rc = ctl_init(t, c, 0.0, 1);
    hipStream_t s = t->stream;
    if (!rc && (hipMemcpyAsync(drp, row_ptr, sizeof(int) * (n + 1), hipMemcpyHostToDevice, s) != hipSuccess ||
                hipMemcpyAsync(dci, col, sizeof(int) * nnz, hipMemcpyHostToDevice, s) != hipSuccess ||
                hipMemcpyAsync(dv, val, sizeof(double) * nnz, hipMemcpyHostToDevice, s) != hipSuccess))
        rc = BKH_DEVICE_ERROR;
    if (!rc) {
        hipLaunchKernelGGL(k_spnorm, dim3(G), dim3(256), 0, s, n, (const int *)drp, (const int *)dci,
                           (const double *)dv, symmetric, part);
        launch_finalize<1>(s, part, G, c, StoreFin());
        if (hipGetLastError() != hipSuccess) rc = BKH_DEVICE_ERROR;
    }
    if (!rc) rc = ctl_fetch(t, c);
    if (!rc) *out = sqrt(t->ctl_host->s[0]);
    hipFree(mem);
    return rc;
}
