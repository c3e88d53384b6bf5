// bkh_host.hpp -- host-side internals of the device layer (C++ inside the .hip
// translation units; the public surface is the extern "C" API in bksp_hip.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <functional>
#include <memory>
#include <vector>
#include <utility>
#include "bkh_core.hpp"
#include <atomic>

// Every device allocation of the library goes through bkh_dmalloc, so a test can make
// the n-th next one fail (bkh_debug_fail_alloc) and check that each caller turns the
// failure into an error return.
extern std::atomic<int> g_bkh_fail_alloc;
template <typename T>
inline hipError_t bkh_dmalloc(T **p, size_t bytes) {
    if (g_bkh_fail_alloc.load(std::memory_order_relaxed) > 0 && g_bkh_fail_alloc.fetch_sub(1) == 1) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc((void **)p, bytes);
}
#include "../../../include/bksp_hip.h"

// Records a message retrievable with bkh_last_error() (thread-local).
void bkh_set_error(const char *fmt, ...);

#define HIP_OK(expr)                                                              \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) {                                                   \
            bkh_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),  \
                          __FILE__, __LINE__);                                    \
            return BKH_DEVICE_ERROR;                                              \
        }                                                                         \
    } while (0)

#define HIP_OK_NULL(expr)                                                         \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) {                                                   \
            bkh_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),  \
                          __FILE__, __LINE__);                                    \
            return nullptr;                                                       \
        }                                                                         \
    } while (0)

#define LAUNCH_OK()                                                               \
    do {                                                                          \
        hipError_t e_ = hipGetLastError();                                        \
        if (e_ != hipSuccess) {                                                   \
            bkh_set_error("kernel launch failed: %s (%s:%d)",                     \
                          hipGetErrorString(e_), __FILE__, __LINE__);             \
            return BKH_DEVICE_ERROR;                                              \
        }                                                                         \
    } while (0)

// CSR-stream row blocks: a block of <= SPMV_ROWS rows whose nonzeros fit in
// SPMV_CAP LDS slots is streamed into LDS with coalesced loads; a single row with
// more nonzeros than SPMV_CAP forms a block of its own (workgroup reduction).
#define SPMV_BS 256
#define SPMV_ROWS 256
#define SPMV_CAP 2048
// Wave-stream variant: each 64-lane wave owns <= 64 rows / <= SPMV_WCAP nonzeros
// and needs no workgroup barrier; 4 waves per workgroup.
#define SPMV_WROWS 64
#define SPMV_WCAP 512
// SELL-64: rows in slices of 64 (one per wavefront); entry k of the slice's row l
// sits at slice_offset + 64 k + l, padded to the slice's longest row, so a lane owns
// a row and every load of a step is one contiguous 64-entry run. Used when padding
// stays below SELL_MAX_PAD of the nonzeros.
#define SELL_MAX_PAD 1.2
enum { SPMV_BLOCK = 0, SPMV_WAVE = 1, SPMV_SELL = 2, SPMV_SELL32 = 3 };
// Slice-diagonal copy (banded matrices, e.g. stencils): the SELL-64 slices with every
// stored position k of a slice carrying one column offset o_k for all 64 rows
// (column = row + o_k), the offsets of a slice the ascending union of its rows'
// column - row values, absent entries stored as value 0 at a column clamped into
// [0, n). No per-entry index stream; built when every row's columns ascend (so the
// ascending offsets keep each row's summation order) and the copy stays below
// DIA_MAX_PAD of the nonzeros. Used by the CG iteration kernels.
#define DIA_MAX_PAD 1.1
// One 64-byte record per slice of the slice-diagonal copy, read with a single scalar
// load by the pipelined block kernels: the slice's first entry, its number of stored
// positions, the position whose offset is 0, whether every row + offset of the slice
// lies in [0, n) (then no gather needs a clamp), and the offsets (past `len`, the last
// one repeated). Built when every slice has 1..DIA_META_POS positions and an offset 0.
#define DIA_META_POS 12
struct DiaSlice {
    int s0, len, cq, inter;
    int o[DIA_META_POS];
};

// Device-resident CSR. Rows are stored in the reference's accumulation order
// (see bkh_matrix_create), so a sequential per-row sum reproduces src/matvec.c.
struct bkh_matrix {
    int n = 0;
    long long nnz = 0;                 // expanded nonzeros
    int *rp = nullptr, *ci = nullptr;  // row_ptr[n+1], col[nnz]
    double *val = nullptr;             // FP64 values
    float *val32 = nullptr;            // FP32 copy (precision extension)
    short *cd16 = nullptr;             // column - first row of the entry's wave block,
                                       // when every such delta fits in 16 bits
    double *diag = nullptr;            // diagonal (for Jacobi)
    int nblk = 0;                      // CSR-stream row blocks (workgroup variant)
    int *blk = nullptr;                // blk[nblk+1] row boundaries
    int nwblk = 0;                     // row blocks of the wave variant
    int *wblk = nullptr;               // wblk[nwblk+1]
    int nsl = 0;                       // SELL-64 slices (0: no SELL copy)
    long long sell_entries = 0;        // stored SELL entries (nonzeros + padding)
    int *sso = nullptr;                // slice offsets in entries, nsl + 1
    double *sval = nullptr;            // SELL values (padding 0)
    float *sval32 = nullptr;           // FP32 copy
    int *scol = nullptr;               // SELL columns (padding: the row itself)
    short *scd = nullptr;              // column - row, when every delta fits in 16 bits
    long long dia_entries = 0;         // slice-diagonal copy (0: none)
    int *dso = nullptr;                // its slice offsets in entries, nsl + 1
    double *dval = nullptr;            // its values (absent entries 0)
    float *dval32 = nullptr;           // FP32 copy
    int *doff = nullptr;               // column offset of stored position k of slice q at dso[q] / 64 + k
    DiaSlice *dmeta = nullptr;         // per-slice records (nullptr: some slice does not qualify)
    int dia_maxpos = 0;                // most stored positions of any slice
    int dia_maxoff = 0;                // largest |column offset| of the copy
    int variant = SPMV_WAVE;           // kernel used by every SpMV on this matrix
    int wrows = SPMV_WROWS, wcap = SPMV_WCAP;   // the wave variant's rows / nonzeros per wavefront
    int spmm_w[17] = {};               // bkh_spmm: columns per pass for s = 1..16 (0: s)
    int max_row = 0;                   // longest row (nonzeros)
    bkh_matrix *T = nullptr;           // A^T in the same format (BiCG)
    int stored_triangle = 0;
};

// Per-thread device context: one stream and a grow-only workspace arena.
struct ThreadCtx {
    hipStream_t stream = nullptr;
    void *ws = nullptr;
    size_t ws_bytes = 0;
    Ctl *ctl_host = nullptr;           // pinned mirror of the control block
    double *red2 = nullptr;            // second-level partial sums (RED2_CAP doubles)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<std::pair<std::vector<uintptr_t>, hipGraphExec_t>> graphs;   // graph cache
    // eager mode: the last few iteration bodies, run as plain launches
    std::vector<std::unique_ptr<std::pair<std::function<void()>, int>>> eager;
};
ThreadCtx *bkh_tctx();                 // nullptr (with error set) if init failed
#define RED2_CAP (1 << 17)
// Returns a device workspace of at least `bytes` (valid until the next call).
void *bkh_workspace(size_t bytes);

// Carves aligned sub-buffers out of one workspace allocation.
struct Carve {
    char *base;
    size_t off = 0;
    explicit Carve(void *b) : base((char *)b) {}
    template <class T> T *take(size_t count) {
        off = (off + 255) & ~(size_t)255;
        T *p = (T *)(base + off);
        off += count * sizeof(T);
        return p;
    }
};
static inline size_t carve_size(size_t bytes) { return ((bytes + 255) & ~(size_t)255) + 256; }

// Grid for streaming vector kernels.
static inline int vec_grid(long long n, int bs, int per_thread = 4) {
    long long g = (n + (long long)bs * per_thread - 1) / ((long long)bs * per_thread);
    if (g < 1) g = 1;
    if (g > 4096) g = 4096;
    return (int)g;
}

// Runs `body(stream)` K times inside a stream capture and returns an executable
// graph, or nullptr on failure.
template <class Body>
hipGraphExec_t capture_graph(hipStream_t s, int K, Body body) {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
    for (int k = 0; k < K; k++) body();
    if (hipStreamEndCapture(s, &g) != hipSuccess || !g) return nullptr;
    if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) ge = nullptr;
    hipGraphDestroy(g);
    return ge;
}

// Cache of instantiated iteration graphs, per thread. A solver's graph depends only
// on the matrix, the vectors and workspace it was captured with and a few launch
// parameters; the key holds all of them plus a global epoch that is advanced
// whenever a matrix is destroyed or re-laid out (so an address reused by another
// matrix never matches). Cached graphs are owned by the cache.
typedef std::vector<uintptr_t> GraphKey;
hipGraphExec_t graph_cache_find(ThreadCtx *t, const GraphKey &key);
void graph_cache_put(ThreadCtx *t, const GraphKey &key, hipGraphExec_t ge);
uintptr_t bkh_graph_epoch();
void bkh_graph_epoch_bump();
// bkh_spmm's pass over s column-major columns in chunks of w (blcg.hip); the autotuner
// times it per width
int spmm_chunked(ThreadCtx *t, const bkh_matrix *A, int s, int w, double alpha, const double *X, double beta,
                 double *Y);

// Eager mode (environment variable BKH_EAGER=1): an iteration body is kept as a
// function and run_graph_until_done issues its launches directly on the stream
// instead of replaying a graph -- the same kernels in the same order, for tools that
// do not follow graph launches. The handle is valid for the rest of the calling
// solve (the body refers to the caller's locals); the last few are kept per thread.
bool bkh_eager();
hipGraphExec_t eager_handle(ThreadCtx *t, std::function<void()> body, int K);

template <class Body>
hipGraphExec_t capture_cached(ThreadCtx *t, GraphKey key, hipStream_t s, int K, Body body) {
    if (bkh_eager()) return eager_handle(t, std::function<void()>(body), K);
    key.push_back(bkh_graph_epoch());
    key.push_back((uintptr_t)t->ws);
    key.push_back((uintptr_t)t->ws_bytes);
    key.push_back((uintptr_t)K);
    if (hipGraphExec_t ge = graph_cache_find(t, key)) return ge;
    hipGraphExec_t ge = capture_graph(s, K, body);
    if (ge) graph_cache_put(t, key, ge);
    return ge;
}

// Replays `ge` until the control block says done. Keeps one graph in flight
// ahead of the host check. Returns 0 or BKH_DEVICE_ERROR.
int run_graph_until_done(ThreadCtx *t, hipGraphExec_t ge, Ctl *ctl_dev, int max_launches);

// Initialises a control block on the device (async, on the thread's stream);
// it0 is the loop index of the first iteration (CR's loop starts at 1).
int ctl_init(ThreadCtx *t, Ctl *ctl_dev, double tol, int max_it, int it0 = 0);
// Copies the control block to t->ctl_host and synchronises.
int ctl_fetch(ThreadCtx *t, const Ctl *ctl_dev);

// Maps a finished control block to the bksp return convention.
int ctl_result(const Ctl *c);
