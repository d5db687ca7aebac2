// api.hip -- the C ABI of librbffd (declared in include/rbffd.h): context,
// device workspace, the generation pipeline (kNN -> nearest X of every Y ->
// local solves -> FP64 re-solve of unconverged rows), CSR -> CSC and SpMV.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <climits>
#include <condition_variable>
#include <cstdlib>
#include <exception>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include <sched.h>

#include "internal.hpp"
#include "rbffd.h"

using rbffd::Error;

struct rbffd_buf {
    void *p = nullptr;
    size_t bytes = 0;
};

// Workspace slots of one neighbour search and the per-set buffers of a batch.
enum ArenaSlot {
    A_PART, A_BOX, A_GRID, A_COUNTS, A_START, A_CELL_OF, A_SLOT, A_CELL_XY, A_CELL_IDX, A_CELL_GRP,
    A_SCAN_TMP, A_STENCILS, A_STATUS, A_REDO, A_REDO_TAU, A_REDO2, A_GEN_SCRATCH, A_NSLOT
};

struct rbffd_arena {
    rbffd_buf b[A_NSLOT];
};

constexpr int kLanes = 4;  // parallel branches of a captured batch

// A few persistent host threads that run one job at a time (the host side of
// the staged device-to-host copies).
struct rbffd_host_pool {
    std::vector<std::thread> threads;
    std::mutex mu;
    std::condition_variable wake, done;
    std::function<void(int, int)> job;
    long generation = 0;
    int pending = 0;
    bool stop = false;

    explicit rbffd_host_pool(int n)
    {
        for (int t = 0; t < n; ++t) threads.emplace_back([this, t, n] { loop(t, n); });
    }
    ~rbffd_host_pool()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        wake.notify_all();
        for (auto &th : threads) th.join();
    }
    void loop(int t, int n)
    {
        long seen = 0;
        for (;;) {
            std::function<void(int, int)> f;
            {
                std::unique_lock<std::mutex> lk(mu);
                wake.wait(lk, [&] { return stop || generation != seen; });
                if (stop) return;
                seen = generation;
                f = job;
            }
            f(t, n);
            {
                std::lock_guard<std::mutex> lk(mu);
                if (--pending == 0) done.notify_all();
            }
        }
    }
    // run f(thread, threads) on every thread and wait for all of them
    void run(std::function<void(int, int)> f)
    {
        std::unique_lock<std::mutex> lk(mu);
        job = std::move(f);
        pending = (int)threads.size();
        ++generation;
        wake.notify_all();
        done.wait(lk, [&] { return pending == 0; });
    }
};

// Workspace slots of the context (one grow-only device buffer each).
enum Slot {
    S_BBOX_PART, S_BBOX, S_FLAG, S_XGROUP, S_QGROUP, S_STENCILS, S_Y2X, S_STATUS, S_ROWS, S_CSC_COUNTS,
    S_CSC_PTR, S_CSC_FILL, S_CSC_PERM, S_CSC_ROW, S_H_X, S_H_Y, S_H_COL, S_H_VALS, S_H_CVALS,
    S_H_PTR64, S_H_ROW64, S_SCAN_TMP2, S_BATCH_FLAGS, S_MORTON_KEY, S_MORTON_IDX, S_CSC_PERM_IN, S_LSQ_ROWOF, S_LSQ_TCOL, S_LSQ_TPTR, S_LSQ_F64,
    S_LSQ_COUNT, S_LSQ_DENSE, S_LSQ_INFO, S_CLOSURE_T,
    S_CLOSURE_PART, S_GRP_COUNTS, S_GRP_YSTART, S_GRP_FILL, S_GRP_NTASK, S_GRP_TSTART, S_GRP_YLIST, S_GRP_TASKS,
    S_GRP_TMP, S_GEN_SCRATCH, S_NSLOT
};

struct rbffd_context {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    using Buf = rbffd_buf;
    Buf bufs[S_NSLOT];
    hipEvent_t ev[8] = {};
    hipEvent_t tune_ev[2] = {};
    // autotuner: (n, P, right-hand sides) -> lanes per stencil of the FP64 kernel
    std::map<std::tuple<int, int, int>, int> lu_group;
    // (n, P, right-hand sides) -> whether Y rows sharing a stencil solve grouped
    std::map<std::tuple<int, int, int, bool>, bool> group_rows;
    int last_group = 0;  // lanes per stencil of the last solve launch
    // neighbour-search workspace of the single-set calls, and one per batch branch
    rbffd_arena knn_arena;
    rbffd_arena lane_arena[kLanes];
    hipStream_t lane_stream[kLanes] = {};
    hipEvent_t fork_ev = nullptr, join_ev[kLanes] = {};
    // the last captured batch: its arguments, the workspace buffers it
    // recorded, and its executable graph
    std::vector<int64_t> batch_key;
    std::vector<const void *> batch_bufs;
    hipGraph_t batch_graph = nullptr;
    hipGraphExec_t batch_exec = nullptr;
    // staged device-to-host copies: two pinned buffers and host copy threads
    void *stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    std::unique_ptr<rbffd_host_pool> pool;
    // rocSOLVER handle of the direct sparse solve (created on first use)
    void *solver = nullptr;
    // this device's copy of the static workspace of the no-allocation entry point
    char *noalloc = nullptr;
};

// Device workspace of rbffd_weights_device_noalloc: a fixed array in every
// GPU's memory, held until the process ends.  Host code never passes this
// name to a kernel or a copy (on the host it only stands for the symbol): it
// uses the device address hipGetSymbolAddress returns for the current device,
// which rbffd_create looks up once per device.
constexpr size_t kNoAllocBytes = 64u << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

thread_local std::string g_last_error;

// scratch's device address, per device (looked up once, under the lock)
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices] = {};

char *scratch_address(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw Error{RBFFD_ERR_RUNTIME, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
        g_scratch_ptr[device] = (char *)ptr;
    }
    return g_scratch_ptr[device];
}


template <class F>
int guarded(F &&f)
{
    try {
        f();
        return RBFFD_OK;
    } catch (const Error &e) {
        g_last_error = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return RBFFD_ERR_RUNTIME;
    } catch (...) {
        g_last_error = "unknown failure";
        return RBFFD_ERR_RUNTIME;
    }
}

// Grow-only device buffer.  With may_alloc false (while a graph is being
// captured) a buffer that is too small is an error instead of an allocation.
void *grow(rbffd_buf &b, size_t bytes, bool may_alloc = true)
{
    if (bytes == 0) bytes = 16;
    if (b.bytes < bytes) {
        if (!may_alloc) throw Error{RBFFD_ERR_RUNTIME, "batch workspace changed between runs"};
        if (b.p) RBFFD_HIP(hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
        size_t want = bytes + bytes / 8;
        RBFFD_HIP(hipMalloc(&b.p, want));
        b.bytes = want;
    }
    return b.p;
}

void *ws(rbffd_context *c, Slot slot, size_t bytes) { return grow(c->bufs[slot], bytes); }

// Where a call takes its device buffers: the context's grow-only slots, or,
// for the no-allocation entry point, consecutive pieces of the per-device
// static array (which must hold them all; nothing is allocated).
struct Workspace {
    rbffd_context *c;
    char *arena = nullptr;
    size_t cap = 0, used = 0;
    bool may_alloc = true;

    void *take(size_t bytes)
    {
        const size_t b = ((bytes ? bytes : 16) + 255) & ~(size_t)255;
        if (b > cap - used)
            throw Error{RBFFD_ERR_ARG, "the call needs more device workspace than the " +
                                           std::to_string(cap >> 20) + " MiB static array holds"};
        void *p = arena + used;
        used += b;
        return p;
    }
    void *get(Slot slot, size_t bytes) { return arena ? take(bytes) : grow(c->bufs[slot], bytes, may_alloc); }
    void *get(rbffd_buf &b, size_t bytes) { return arena ? take(bytes) : grow(b, bytes, may_alloc); }
    // bytes still available (unbounded for the grow-only slots)
    size_t room() const { return arena ? cap - used : (size_t)-1; }
};

void check_ctx(rbffd_context *c)
{
    if (!c) throw Error{RBFFD_ERR_ARG, "null context"};
    RBFFD_HIP(hipSetDevice(c->device));
}

// ------------------------------------------------------------ small kernels
__global__ void count_negative_kernel(const int *v, long n, int *flag)
{
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n && v[i] < 0) atomicAdd(flag, 1);
}

// Largest value: a block maximum, then one atomic per block.
__global__ __launch_bounds__(256) void max_kernel(const int *v, long n, int *out)
{
    __shared__ int sm[256];
    int m = INT_MIN;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        m = max(m, v[i]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] = max(sm[threadIdx.x], sm[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, sm[0]);
}

__global__ void status_count_kernel(const int *status, long n, int *counts)
{
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) {
        int s = status[i];
        if (s > 0 && s < 4) atomicAdd(&counts[s], 1);
    }
}

__global__ void collect_kernel(const int *status, long n, int code, int *list, int *counter)
{
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n && status[i] == code) list[atomicAdd(counter, 1)] = (int)i;
}

inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

int read_flag(rbffd_context *c, const int *dflag)
{
    int h = 0;
    RBFFD_HIP(hipMemcpyAsync(&h, dflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    RBFFD_HIP(hipStreamSynchronize(c->stream));
    return h;
}

// Device-to-host copy into pageable caller memory, pipelined: chunks go over
// PCIe into two pinned buffers in turn while the host threads copy the
// previous chunk into place.  Synchronous.
constexpr size_t kStageBytes = 32u << 20;

// Host threads that copy staged chunks into the caller's arrays: the job's CPU
// share (RBFFD_HOST_THREADS, else OMP_NUM_THREADS, else the CPUs this process
// may run on), at most 32.  Writes into pageable memory bound the copy-out:
// 1.05 GB took 45 ms with 8 threads and 37 ms with 16 on a 16-CPU share.
int host_copy_threads()
{
    for (const char *name : {"RBFFD_HOST_THREADS", "OMP_NUM_THREADS"}) {
        const char *e = std::getenv(name);
        if (e && std::atoi(e) > 0) return std::min(32, std::atoi(e));
    }
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) return std::max(1, std::min(32, CPU_COUNT(&set)));
    return (int)std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
}

void d2h_staged(rbffd_context *c, void *dst, const void *src, size_t bytes)
{
    hipStream_t s = c->stream;
    if (bytes < 2 * kStageBytes) {
        RBFFD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        RBFFD_HIP(hipStreamSynchronize(s));
        return;
    }
    for (int b = 0; b < 2; ++b) {
        if (!c->stage[b]) RBFFD_HIP(hipHostMalloc(&c->stage[b], kStageBytes, hipHostMallocDefault));
        if (!c->stage_ev[b]) RBFFD_HIP(hipEventCreateWithFlags(&c->stage_ev[b], hipEventDisableTiming));
    }
    if (!c->pool) c->pool = std::make_unique<rbffd_host_pool>(host_copy_threads());
    const size_t nchunk = (bytes + kStageBytes - 1) / kStageBytes;
    auto enqueue = [&](size_t i) {
        const size_t off = i * kStageBytes, len = std::min(kStageBytes, bytes - off);
        RBFFD_HIP(hipMemcpyAsync(c->stage[i & 1], (const char *)src + off, len, hipMemcpyDeviceToHost, s));
        RBFFD_HIP(hipEventRecord(c->stage_ev[i & 1], s));
    };
    enqueue(0);
    for (size_t i = 0; i < nchunk; ++i) {
        // chunk i+1 reuses the buffer whose host copy finished in the previous round
        if (i + 1 < nchunk) enqueue(i + 1);
        RBFFD_HIP(hipEventSynchronize(c->stage_ev[i & 1]));
        const size_t off = i * kStageBytes, len = std::min(kStageBytes, bytes - off);
        const char *from = (const char *)c->stage[i & 1];
        char *to = (char *)dst + off;
        c->pool->run([&](int t, int n) {
            const size_t part = (len + n - 1) / n, a = std::min(len, part * t), b = std::min(len, a + part);
            if (b > a) memcpy(to + a, from + a, b - a);
        });
    }
}

// ------------------------------------------------------------ kNN
// Enqueue one neighbour search on stream s: bounding box, grid parameters
// (on the device, so nothing is read back), counting sort of X by cell, then
// the query kernel.  flag[0] |= 1 when X is not finite; flag[1] counts
// result slots left at -1 (a query seeing fewer than k nodes).  Neither
// synchronises nor (with may_alloc false) allocates, so it can be captured.
void knn_enqueue(rbffd_arena &A, bool may_alloc, const double *X, long N, const double *Q, long M, int k,
                 const int *xgroup, const int *qgroup, int *idx, double *d2, int *flag, hipStream_t s)
{
    const int nb = (int)std::min<long>(1024, (N + 255) / 256);
    double *part = (double *)grow(A.b[A_PART], sizeof(double) * 5 * nb, may_alloc);
    double *box = (double *)grow(A.b[A_BOX], sizeof(double) * 5, may_alloc);
    auto *grid = (rbffd::GridParams *)grow(A.b[A_GRID], sizeof(rbffd::GridParams), may_alloc);
    const long ncell = rbffd::grid_max_cells(N);
    int *counts = (int *)grow(A.b[A_COUNTS], sizeof(int) * (ncell + 1), may_alloc);
    int *start = (int *)grow(A.b[A_START], sizeof(int) * (ncell + 1), may_alloc);
    int *cell_of = (int *)grow(A.b[A_CELL_OF], sizeof(int) * N, may_alloc);
    int *slot = (int *)grow(A.b[A_SLOT], sizeof(int) * N, may_alloc);
    double *cxy = (double *)grow(A.b[A_CELL_XY], sizeof(double) * 2 * N, may_alloc);
    int *cidx = (int *)grow(A.b[A_CELL_IDX], sizeof(int) * N, may_alloc);
    int *cgrp = xgroup ? (int *)grow(A.b[A_CELL_GRP], sizeof(int) * N, may_alloc) : nullptr;
    const size_t tb = rbffd::scan_tmp_bytes(ncell + 1);
    void *tmp = grow(A.b[A_SCAN_TMP], tb, may_alloc);

    rbffd::launch_bbox(X, N, box, part, nb, s);
    rbffd::launch_grid_params(box, N, k, grid, flag, s);
    rbffd::launch_fill_i32(counts, ncell + 1, 0, s);
    rbffd::launch_cell_count(X, N, grid, cell_of, counts, slot, s);
    rbffd::launch_exclusive_scan(counts, start, ncell + 1, tmp, tb, s);
    rbffd::launch_cell_scatter(X, xgroup, N, cell_of, start, slot, cxy, cidx, cgrp, s);
    // X queries run in cell order so a wavefront walks neighbouring cells
    const int *qorder = (Q == X && M == N) ? cidx : nullptr;
    int *redo = (int *)grow(A.b[A_REDO], sizeof(int) * (M + 1), may_alloc);
    rbffd::KnnRedo rw;
    rw.tau = (double *)grow(A.b[A_REDO_TAU], sizeof(double) * M, may_alloc);
    int *redo2 = (int *)grow(A.b[A_REDO2], sizeof(int) * (M + 1), may_alloc);
    rw.count2 = redo2;
    rw.list2 = redo2 + 1;
    rbffd::launch_knn_query(grid, start, cxy, cidx, cgrp, N, Q, qgroup, qorder, M, k, idx, d2, flag + 1, redo + 1,
                            redo, rw, s);
    RBFFD_HIP(hipGetLastError());
}

void knn_impl(rbffd_context *c, const double *X, long N, const double *Q, long M, int k,
              const int *xgroup, const int *qgroup, int *idx, double *d2)
{
    if (N < 1) throw Error{RBFFD_ERR_ARG, "knn: no points"};
    if (k < 1 || k > N) throw Error{RBFFD_ERR_ARG, "knn: k must be in 1..number of points"};
    if ((xgroup == nullptr) != (qgroup == nullptr))
        throw Error{RBFFD_ERR_ARG, "knn: xgroup and qgroup go together"};
    if (M <= 0) return;
    hipStream_t s = c->stream;
    int *flag = (int *)ws(c, S_FLAG, sizeof(int) * 8);
    rbffd::launch_fill_i32(flag, 2, 0, s);
    knn_enqueue(c->knn_arena, true, X, N, Q, M, k, xgroup, qgroup, idx, d2, flag, s);
    int h[2] = {0, 0};
    RBFFD_HIP(hipMemcpyAsync(h, flag, sizeof(h), hipMemcpyDeviceToHost, s));
    RBFFD_HIP(hipStreamSynchronize(s));
    if (h[0] & 1) throw Error{RBFFD_ERR_NONFINITE, "knn: non-finite coordinates"};
    if (h[1] != 0) throw Error{RBFFD_ERR_ARG, "knn: a query sees fewer than k nodes"};
}

int n_monomials(int polydeg) { return (polydeg + 1) * (polydeg + 2) / 2; }

void check_ops(int nops, const int *ops, const int *opk)
{
    if (nops < 1 || nops > RBFFD_MAX_OPS) throw Error{RBFFD_ERR_ARG, "nops must be in 1..10"};
    for (int o = 0; o < nops; ++o) {
        if (ops[o] < 0 || ops[o] > RBFFD_OP_LAPK) throw Error{RBFFD_ERR_ARG, "unknown operator kind"};
        if ((ops[o] == RBFFD_OP_XK || ops[o] == RBFFD_OP_YK || ops[o] == RBFFD_OP_LAPK) &&
            (!opk || opk[o] < 0 || opk[o] > 12))
            throw Error{RBFFD_ERR_ARG, "derivative order k must be in 0..12"};
    }
}

// Lanes per stencil for the FP64 kernel.  Where several layouts are compiled
// for a shape, the first call with that shape on a context times each on a
// sample of the batch (the sample's outputs are rewritten by the real launch)
// and keeps the fastest.  RBFFD_LU_GROUP=<lanes> forces a layout.
int choose_group(rbffd_context *c, const rbffd::SolveArgs &a, bool may_tune = true)
{
    int groups[rbffd::kMaxGroups];
    rbffd::solve_groups(a.n, a.P, 0, groups);
    int ncand = 0;
    while (ncand < rbffd::kMaxGroups && groups[ncand] != 0) ++ncand;
    if (const char *env = std::getenv("RBFFD_LU_GROUP")) {
        const int g = std::atoi(env);
        for (int gi = 0; gi < ncand; ++gi)
            if (g == groups[gi]) return g;
    }
    if (ncand <= 1) return groups[0];
    const int rb = a.nops <= 3 ? a.nops : rbffd::kMaxRhs;
    auto key = std::make_tuple(a.n, a.P, rb);
    auto it = c->lu_group.find(key);
    if (it != c->lu_group.end()) return it->second;
    if (a.ntasks < 4096 || !may_tune) return groups[0];  // too small to time (or capturing): the default
    for (auto &e : c->tune_ev)
        if (!e) RBFFD_HIP(hipEventCreate(&e));
    rbffd::SolveArgs t = a;
    t.ntasks = std::min<long>(a.ntasks, 1L << 16);
    t.status_accumulate = 0;
    float best = 1e30f;
    int pick = groups[0];
    for (int rep = 0; rep < 2; ++rep) {  // the first pass warms up every kernel
        for (int gi = 0; gi < ncand; ++gi) {
            t.group = groups[gi];
            RBFFD_HIP(hipEventRecord(c->tune_ev[0], c->stream));
            rbffd::launch_solve(t, 0, c->stream);
            RBFFD_HIP(hipEventRecord(c->tune_ev[1], c->stream));
            RBFFD_HIP(hipEventSynchronize(c->tune_ev[1]));
            float ms = 0;
            RBFFD_HIP(hipEventElapsedTime(&ms, c->tune_ev[0], c->tune_ev[1]));
            if (rep == 1 && ms < best) { best = ms; pick = groups[gi]; }
        }
    }
    c->lu_group[key] = pick;
    return pick;
}

// The precision RBFFD_PREC_AUTO resolves to.  FP64: on MI355X the FP32 bulk
// solve issues at the FP64 rate (packed FP32 is not used), so the mixed path
// only adds refinement work.
int auto_precision() { return RBFFD_PREC_FP64; }

void check_weights_args(int n, int p, int polydeg, int method, int nops, const int *ops, const int *opk,
                        int precision)
{
    if (method < 1 || method > 3) throw Error{RBFFD_ERR_ARG, "method must be 1, 2 or 3"};
    if (p < 1) throw Error{RBFFD_ERR_ARG, "PHS power must be >= 1"};
    if (polydeg < 0 || polydeg > 30) throw Error{RBFFD_ERR_ARG, "polynomial degree must be in 0..30"};
    if (n < 1) throw Error{RBFFD_ERR_ARG, "stencil size must be >= 1"};
    check_ops(nops, ops, opk);
    if (precision != RBFFD_PREC_FP64 && precision != RBFFD_PREC_MIXED && precision != RBFFD_PREC_AUTO)
        throw Error{RBFFD_ERR_ARG, "unknown precision"};
    const int P = n_monomials(polydeg);
    if (!rbffd::solve_supported(n, P))
        throw Error{RBFFD_ERR_UNSUPPORTED, "stencil size n=" + std::to_string(n) + " with polydeg=" +
                                               std::to_string(polydeg) +
                                               ": n must be in 1..128 and n + P at most 512"};
}

// Y rows that share an X node's stencil are solved together (one
// factorisation per task) unless RBFFD_GROUP_ROWS=0.
bool grouping_enabled()
{
    const char *env = std::getenv("RBFFD_GROUP_ROWS");
    return !(env && env[0] == '0');
}

// Grouped tasks of a shape with a specialised kernel run in that kernel
// (its one-slot layout) unless RBFFD_TUNED_GROUPS=0 sends them to the generic one.
bool tuned_groups(int n, int P, int precision)
{
    const char *env = std::getenv("RBFFD_TUNED_GROUPS");
    return precision == RBFFD_PREC_FP64 && rbffd::solve_tuned(n, P, 0) && !(env && env[0] == '0');
}

// Output rows per grouped task of the specialised kernel for nops operators.
// once: every row of an X node in one task (one factorisation per X node, as
// the reference) while that fits the widest right-hand-side block.  Else the
// narrowest block that holds the mean rows per X node (narrow rows cost less
// per step than the factorisations they add: 4M rows over 1M X nodes at
// n = 31 take 17.3 ms in tasks of 4 rows, 19.5 ms in tasks of 8, 24.2 ms in
// tasks of 2).  Rows past ky go to further tasks.
int lu_rows_per_task(int nops, int maxrows, double mean_rows, bool once)
{
    if (nops < 1 || maxrows <= 1) return 1;
    int ky = rbffd::kLuGrpMaxRhs / nops;
    if (!once) {
        const int want = (int)(mean_rows + 0.999);
        for (int b = 6; b <= rbffd::kLuGrpMaxRhs; b *= 2) {
            if (b / nops >= want) {
                ky = b / nops;
                break;
            }
        }
    }
    ky = std::min(ky, maxrows);
    return ky < 1 ? 1 : ky;
}

// Largest device-side workspace of the solves of systems above 128 equations.
constexpr long kGenScratchBytes = 1L << 30;
// Rows from which the choice between grouped and per-row solves is timed.
constexpr long kGroupTuneRows = 1L << 16;

// Enqueue the local solves on stream s (the autotuner, when allowed, times on
// the context stream).  Y rows sharing an X node's stencil (y2x given, FP64,
// no row list) are grouped into tasks solved by one factorisation each; that
// reads the largest group size back (a synchronisation), so a captured batch
// (y2x NULL) never groups.  scratch: the workspace of systems above 128
// equations (one per concurrent stream).  Returns the number of
// factorisations (tasks) of the first operator chunk; with grouping it is
// known on the device only, and *ntasks_dev (if not NULL) receives where.
long solves_enqueue(rbffd_context *c, Workspace &w, long N, const double *X, const double *Y, long M,
                    const int *stencils, int n, const int *y2x, const int *rows, long nrows, int p, int polydeg,
                    int method, int nops, const int *ops, const int *opk, int precision, int *colidx,
                    double *const *vals, int *status, hipStream_t s, bool may_tune, rbffd_buf &scratch,
                    const int **ntasks_dev = nullptr)
{
    const int P = n_monomials(polydeg);
    const int cmax = std::min(rbffd::kMaxRhs, nops);
    if (ntasks_dev) *ntasks_dev = nullptr;
    // ---- group the Y rows by their X node
    int ky = 1;
    int *counts = nullptr, *fill = nullptr;
    bool grp_tuned = tuned_groups(n, P, precision);
    if (y2x && !rows && precision == RBFFD_PREC_FP64 && M > 1 && grouping_enabled()) {
        counts = (int *)w.get(S_GRP_COUNTS, sizeof(int) * (N + 1));
        fill = (int *)w.get(S_GRP_FILL, sizeof(int) * N);
        int *flag = (int *)w.get(S_FLAG, sizeof(int) * 8);
        rbffd::launch_group_count(y2x, M, N, counts, fill, s);
        rbffd::launch_fill_i32(flag, 1, 0, s);
        max_kernel<<<(unsigned)std::min<long>(2048, nblk(N)), 256, 0, s>>>(counts, N, flag);
        RBFFD_HIP(hipGetLastError());
        const int maxrows = read_flag(c, flag);
        // calls below the tuning size keep one factorisation per X node: in
        // the specialised kernel while an X node's rows fit its 24-column
        // block, else in the generic kernel (blocks up to 72 columns)
        const bool small = M < kGroupTuneRows;
        if (grp_tuned && maxrows * cmax > rbffd::kLuGrpMaxRhs && small) grp_tuned = false;
        ky = grp_tuned ? lu_rows_per_task(cmax, maxrows, (double)M / (double)N, small)
                       : rbffd::gen_rows_per_task(n, P, cmax, maxrows, (double)M / (double)N);
    }
    const int *grp_x = nullptr, *grp_start = nullptr, *grp_count = nullptr, *ylist = nullptr, *ntd = nullptr;
    if (ky > 1) {
        int *ystart = (int *)w.get(S_GRP_YSTART, sizeof(int) * (N + 1));
        int *ntask = (int *)w.get(S_GRP_NTASK, sizeof(int) * (N + 1));
        int *tstart = (int *)w.get(S_GRP_TSTART, sizeof(int) * (N + 1));
        int *yl = (int *)w.get(S_GRP_YLIST, sizeof(int) * M);
        int *tasks = (int *)w.get(S_GRP_TASKS, sizeof(int) * 3 * M);
        const size_t tb = rbffd::group_tmp_bytes(N);
        void *tmp = w.get(S_GRP_TMP, tb);
        rbffd::launch_group_build(y2x, M, N, ky, counts, ystart, fill, ntask, tstart, yl, tasks, tasks + M,
                                  tasks + 2 * M, tmp, tb, s);
        RBFFD_HIP(hipGetLastError());
        grp_x = tasks;
        grp_start = tasks + M;
        grp_count = tasks + 2 * M;
        ylist = yl;
        ntd = tstart + N;
        if (ntasks_dev) *ntasks_dev = ntd;
    }
    // ---- workspace of the systems above 128 equations
    const long ntask = rows ? nrows : M;
    const long mem_ld = rbffd::gen_mem_ld(n, P);
    double *mem = nullptr;
    long mem_tasks = 0;
    if (mem_ld > 0 && ntask > 0) {
        const long per = (long)sizeof(double) * mem_ld * (long)(n + P + cmax);
        const long room = (long)std::min<size_t>(w.room(), (size_t)kGenScratchBytes);
        mem_tasks = std::min(ntask, room / per);
        if (mem_tasks < 1) mem_tasks = 1;  // w.get reports a static array too small for one system
        mem = (double *)w.get(scratch, per * mem_tasks);
    }
    // every operator chunk, with tasks of kyv rows (1: one row per task)
    auto launch_all = [&](int kyv) {
        const bool grouped = kyv > 1;
        const bool tuned = !grouped && rbffd::solve_tuned(n, P, precision);
        const bool grp_tuned_now = grouped && grp_tuned;
        // at most kMaxRhs right-hand sides per launch; more operators run in chunks
        for (int o0 = 0; o0 < nops; o0 += rbffd::kMaxRhs) {
            const int cnt = std::min(rbffd::kMaxRhs, nops - o0);
            rbffd::SolveArgs a{};
            a.X = X;
            a.Y = Y;
            a.stencils = stencils;
            a.y2x = y2x;
            a.rows = rows;
            a.ntasks = ntask;
            a.n = n;
            a.P = P;
            a.m = p;
            a.method = method == 3 ? 3 : 1;
            a.nops = cnt;
            for (int o = 0; o < cnt; ++o) {
                a.op_kind[o] = ops[o0 + o];
                a.op_k[o] = opk ? opk[o0 + o] : 0;
                a.vals[o] = vals[o0 + o];
            }
            // right-hand-side columns of the generic kernel: row-major over
            // (output row of the task, operator)
            a.gen_cols = kyv * cnt;
            for (int cc = 0; cc < a.gen_cols; ++cc) {
                const int o = cc % cnt;
                a.col_sub[cc] = (int8_t)(cc / cnt);
                a.col_op[cc] = (int8_t)o;
                a.col_kind[cc] = (int8_t)a.op_kind[o];
                a.col_k[cc] = (int8_t)a.op_k[o];
            }
            if (grouped) {
                a.grp_x = grp_x;
                a.grp_start = grp_start;
                a.grp_count = grp_count;
                a.ylist = ylist;
                a.ntasks_dev = ntd;
                a.grp_tuned = grp_tuned_now ? 1 : 0;
            }
            a.scratch = mem;
            a.mem_ld = mem_ld;
            a.mem_tasks = mem_tasks;
            a.colidx = o0 == 0 ? colidx : nullptr;
            a.status = status;
            a.status_accumulate = o0 > 0;
            a.max_iter = 8;
            a.tol = 4e-16;
            if (tuned && precision == RBFFD_PREC_FP64) {
                a.group = choose_group(c, a, may_tune);
            } else if (tuned) {
                int g[rbffd::kMaxGroups];
                rbffd::solve_groups(n, P, precision, g);
                a.group = g[0];
            } else if (grp_tuned_now) {
                a.group = rbffd::lu_grp_lanes(n);
            } else {
                a.group = rbffd::gen_lanes_used(n, P, a.gen_cols);
            }
            if (o0 == 0) c->last_group = a.group;
            rbffd::launch_solve(a, precision, s);
            RBFFD_HIP(hipGetLastError());
        }
    };
    // A shape with a specialised kernel solves grouped tasks in that kernel's
    // one-slot layout (the generic kernel with RBFFD_TUNED_GROUPS=0); with
    // few rows per X node one solve per row in the default layout can still
    // be faster.  The first large call per (shape, operators, kernel) times
    // both and keeps the faster; smaller calls group.  RBFFD_GROUP_ROWS=1
    // forces grouping.
    bool group = ky > 1;
    if (group && rbffd::solve_tuned(n, P, precision)) {
        const char *env = std::getenv("RBFFD_GROUP_ROWS");
        const bool forced = env && env[0] == '1';
        auto key = std::make_tuple(n, P, cmax, grp_tuned);
        auto it = c->group_rows.find(key);
        if (forced) {
            group = true;
        } else if (it != c->group_rows.end()) {
            group = it->second;
        } else if (may_tune && M >= kGroupTuneRows && s == c->stream) {
            for (auto &e : c->tune_ev)
                if (!e) RBFFD_HIP(hipEventCreate(&e));
            float ms[2] = {0, 0};
            for (int v = 0; v < 2; ++v) {
                RBFFD_HIP(hipEventRecord(c->tune_ev[0], s));
                launch_all(v == 0 ? ky : 1);
                RBFFD_HIP(hipEventRecord(c->tune_ev[1], s));
                RBFFD_HIP(hipEventSynchronize(c->tune_ev[1]));
                RBFFD_HIP(hipEventElapsedTime(&ms[v], c->tune_ev[0], c->tune_ev[1]));
            }
            group = ms[0] < ms[1];
            c->group_rows[key] = group;
            if (!group) {  // the last run was the faster one
                if (ntasks_dev) *ntasks_dev = nullptr;
                return ntask;
            }
        }
    }
    launch_all(group ? ky : 1);
    if (!group && ntasks_dev) *ntasks_dev = nullptr;
    return ntask;
}

long weights_impl(rbffd_context *c, const double *X, long N, const double *Y, long M, const int *stencils,
                  int n, const int *y2x, const int *rows, long nrows, int p, int polydeg, int method,
                  int nops, const int *ops, const int *opk, int precision, int *colidx,
                  double *const *vals, int *status, const int **ntasks_dev = nullptr, Workspace *wsp = nullptr)
{
    if (precision == RBFFD_PREC_AUTO) precision = auto_precision();
    check_weights_args(n, p, polydeg, method, nops, ops, opk, precision);
    Workspace own{c};
    Workspace &w = wsp ? *wsp : own;
    // the static array admits no autotuning (its events are allocations)
    return solves_enqueue(c, w, N, X, Y, M, stencils, n, y2x, rows, nrows, p, polydeg, method, nops, ops, opk,
                          precision, colidx, vals, status, c->stream, wsp == nullptr, c->bufs[S_GEN_SCRATCH],
                          ntasks_dev);
}

// group ids that restate calculateneighbors' masks (src/calculateneighbors.jl:16-42, 83-87)
void boundary_groups(const rbffd_boundary *b, long N, std::vector<int> &xg, std::vector<int> &qg)
{
    xg.assign(N, -2);
    qg.assign(N, -1);
    std::vector<char> seen(N, 0);
    const long base = b->index_base;
    auto at = [&](int64_t v) -> long {
        long i = (long)v - base;
        if (i < 0 || i >= N) throw Error{RBFFD_ERR_ARG, "boundary index out of range"};
        return i;
    };
    for (int64_t t = 0; t < b->n_in; ++t) { long i = at(b->in_idx[t]); xg[i] = 0; seen[i] = 1; }
    for (int m = 0; m < b->n_markers; ++m)
        for (int64_t t = b->bc_ptr[m]; t < b->bc_ptr[m + 1]; ++t) { long i = at(b->bc_idx[t]); xg[i] = 0; seen[i] = 1; }
    for (int m = 0; m < b->n_markers; ++m) {
        for (int64_t t = b->ghost_ptr[m]; t < b->ghost_ptr[m + 1]; ++t) {
            long i = at(b->ghost_idx[t]);
            xg[i] = m + 1;
            qg[i] = m + 1;
            seen[i] = 1;
        }
        for (int64_t t = b->bc_ptr[m]; t < b->bc_ptr[m + 1]; ++t) qg[at(b->bc_idx[t])] = m + 1;
    }
    // interior queries search the full tree (src/calculateneighbors.jl:83-87)
    for (int64_t t = 0; t < b->n_in; ++t) qg[at(b->in_idx[t])] = -1;
    for (long i = 0; i < N; ++i)
        if (!seen[i])
            throw Error{RBFFD_ERR_UNDEFINED, "node " + std::to_string(i + base) +
                                                 " is in no index set, so it has no stencil"};
}

void generate_impl(rbffd_context *c, const double *X, long N, const double *Y, long M, int p, int n,
                   int polydeg, const rbffd_boundary *bnd, int nops, const int *ops, const int *opk,
                   int precision, int *colidx, double *const *vals, int *stencils_out, int *y2x_out,
                   rbffd_stats *stats, int *status_codes)
{
    hipStream_t s = c->stream;
    if (precision == RBFFD_PREC_AUTO) precision = auto_precision();
    const bool m3 = (Y == nullptr);
    if (m3) { Y = X; M = N; }
    if (n < 1 || n > N) throw Error{RBFFD_ERR_ARG, "stencil size must be in 1..number of X nodes"};
    if (m3 && bnd) throw Error{RBFFD_ERR_ARG, "boundary sets need Y"};
    if (M == 0) {  // no Y nodes: the reference's loops run zero times
        if (stats) *stats = rbffd_stats{};
        return;
    }
    for (int e = 0; e < 8; ++e)
        if (!c->ev[e]) RBFFD_HIP(hipEventCreate(&c->ev[e]));
    RBFFD_HIP(hipEventRecord(c->ev[0], s));

    int *st = stencils_out ? stencils_out : (int *)ws(c, S_STENCILS, sizeof(int) * N * (size_t)n);
    if (bnd) {
        std::vector<int> xg, qg;
        boundary_groups(bnd, N, xg, qg);
        int *dxg = (int *)ws(c, S_XGROUP, sizeof(int) * N);
        int *dqg = (int *)ws(c, S_QGROUP, sizeof(int) * N);
        RBFFD_HIP(hipMemcpyAsync(dxg, xg.data(), sizeof(int) * N, hipMemcpyHostToDevice, s));
        RBFFD_HIP(hipMemcpyAsync(dqg, qg.data(), sizeof(int) * N, hipMemcpyHostToDevice, s));
        knn_impl(c, X, N, X, N, n, dxg, dqg, st, nullptr);
    } else {
        knn_impl(c, X, N, X, N, n, nullptr, nullptr, st, nullptr);
    }
    const int *y2x = nullptr;
    if (!m3 && Y != X) {
        int *d = y2x_out ? y2x_out : (int *)ws(c, S_Y2X, sizeof(int) * M);
        // nearest X node of every Y node over all of X (src/generate_operator.jl:47,
        // src/calculateneighbors.jl:90-94)
        knn_impl(c, X, N, Y, M, 1, nullptr, nullptr, d, nullptr);
        y2x = d;
    } else if (y2x_out) {
        throw Error{RBFFD_ERR_ARG, "y2x_out needs Y distinct from X"};
    }
    RBFFD_HIP(hipEventRecord(c->ev[1], s));

    int *status = (int *)ws(c, S_STATUS, sizeof(int) * M);
    const int *ntasks_dev = nullptr;
    long nfact = weights_impl(c, X, N, Y, M, st, n, y2x, nullptr, 0, p, polydeg, m3 ? 3 : 1, nops, ops, opk,
                              precision, colidx, vals, status, &ntasks_dev);
    RBFFD_HIP(hipEventRecord(c->ev[2], s));
    if (ntasks_dev) {
        int h = 0;
        RBFFD_HIP(hipMemcpyAsync(&h, ntasks_dev, sizeof(int), hipMemcpyDeviceToHost, s));
        RBFFD_HIP(hipStreamSynchronize(s));
        nfact = h;
    }

    int *flag = (int *)ws(c, S_FLAG, sizeof(int) * 8);
    long nfallback = 0;
    if (precision == RBFFD_PREC_MIXED) {
        int *list = (int *)ws(c, S_ROWS, sizeof(int) * M);
        RBFFD_HIP(hipMemsetAsync(flag, 0, sizeof(int), s));
        collect_kernel<<<nblk(M), 256, 0, s>>>(status, M, 3, list, flag);
        nfallback = read_flag(c, flag);
        if (nfallback > 0)
            weights_impl(c, X, N, Y, M, st, n, y2x, list, nfallback, p, polydeg, m3 ? 3 : 1, nops, ops, opk,
                         RBFFD_PREC_FP64, colidx, vals, status);
    }
    RBFFD_HIP(hipEventRecord(c->ev[3], s));
    RBFFD_HIP(hipMemsetAsync(flag, 0, sizeof(int) * 4, s));
    status_count_kernel<<<nblk(M), 256, 0, s>>>(status, M, flag);
    int hc[4];
    RBFFD_HIP(hipMemcpyAsync(hc, flag, sizeof(hc), hipMemcpyDeviceToHost, s));
    RBFFD_HIP(hipStreamSynchronize(s));
    if (status_codes) {
        RBFFD_HIP(hipMemcpy(status_codes, status, sizeof(int) * M, hipMemcpyDeviceToHost));
    }
    if (stats) {
        stats->rows = M;
        stats->singular = hc[1];
        stats->nonfinite = hc[2];
        stats->refined_fallback = nfallback;
        float t01 = 0, t12 = 0, t23 = 0;
        RBFFD_HIP(hipEventElapsedTime(&t01, c->ev[0], c->ev[1]));
        RBFFD_HIP(hipEventElapsedTime(&t12, c->ev[1], c->ev[2]));
        RBFFD_HIP(hipEventElapsedTime(&t23, c->ev[2], c->ev[3]));
        stats->ms_knn = t01;
        stats->ms_solve = t12;
        stats->ms_fallback = t23;
        stats->lanes_per_stencil = c->last_group;
        stats->factorizations = nfact;
    }
    if (hc[1] > 0)
        throw Error{RBFFD_ERR_SINGULAR, std::to_string(hc[1]) + " stencil systems are exactly singular"};
    if (hc[2] > 0)
        throw Error{RBFFD_ERR_NONFINITE, std::to_string(hc[2]) + " rows have non-finite weights"};
}

void csc_impl(rbffd_context *c, long M, int n, long ncols, const int *colidx, int nops,
              const double *const *csr_vals, int64_t *colptr, int64_t *rowval, double *const *csc_vals,
              int base, int64_t *ncols_used)
{
    hipStream_t s = c->stream;
    const long nnz = M * (long)n;
    if (ncols < 1) throw Error{RBFFD_ERR_ARG, "csc: ncols must be >= 1"};
    // column pointers, the permutation and row indices are int32 inside
    if (nnz > INT_MAX || ncols > INT_MAX || M > INT_MAX)
        throw Error{RBFFD_ERR_ARG, "csc: more than 2^31 - 1 entries, rows or columns"};
    if (nnz == 0) {  // sparse() of empty vectors: no columns referenced
        rbffd::launch_fill_i64(colptr, ncols + 1, base, s);
        if (ncols_used) *ncols_used = 0;
        RBFFD_HIP(hipStreamSynchronize(s));
        return;
    }
    int *flag = (int *)ws(c, S_FLAG, sizeof(int) * 8);
    RBFFD_HIP(hipMemsetAsync(flag, 0, sizeof(int) * 2, s));
    max_kernel<<<(unsigned)std::min<long>(2048, nblk(nnz)), 256, 0, s>>>(colidx, nnz, flag);
    count_negative_kernel<<<nblk(nnz), 256, 0, s>>>(colidx, nnz, flag + 1);
    int hm[2] = {0, 0};
    RBFFD_HIP(hipMemcpyAsync(hm, flag, sizeof(hm), hipMemcpyDeviceToHost, s));
    RBFFD_HIP(hipStreamSynchronize(s));
    if (hm[1] > 0 || hm[0] >= ncols) throw Error{RBFFD_ERR_ARG, "csc: column index out of range"};
    if (ncols_used) *ncols_used = (int64_t)hm[0] + 1;

    int *counts = (int *)ws(c, S_CSC_COUNTS, sizeof(int) * (ncols + 1));
    int *ptr = (int *)ws(c, S_CSC_PTR, sizeof(int) * (ncols + 1));
    int *fill = (int *)ws(c, S_CSC_FILL, sizeof(int) * ncols);
    int *perm = (int *)ws(c, S_CSC_PERM, sizeof(int) * nnz);
    int *row = (int *)ws(c, S_CSC_ROW, sizeof(int) * nnz);
    RBFFD_HIP(hipMemsetAsync(counts, 0, sizeof(int) * (ncols + 1), s));
    RBFFD_HIP(hipMemsetAsync(fill, 0, sizeof(int) * ncols, s));
    rbffd::launch_csc_count(colidx, nnz, counts, s);
    size_t tb = rbffd::scan_tmp_bytes(ncols + 1);
    void *tmp = ws(c, S_SCAN_TMP2, tb);
    rbffd::launch_exclusive_scan(counts, ptr, ncols + 1, tmp, tb, s);
    int *perm_in = (int *)ws(c, S_CSC_PERM_IN, sizeof(int) * nnz);
    rbffd::launch_csc_scatter(colidx, M, n, ptr, fill, perm_in, s);
    rbffd::launch_csc_rank(colidx, M, n, ptr, perm_in, perm, row, s);
    for (int o = 0; o < nops; ++o) rbffd::launch_gather_f64(csr_vals[o], perm, nnz, csc_vals[o], s);
    rbffd::launch_i32_to_i64(ptr, ncols + 1, colptr, base, s);
    rbffd::launch_i32_to_i64(row, nnz, rowval, base, s);
    RBFFD_HIP(hipGetLastError());
    RBFFD_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------ sparse solve
// rows whose rowptr decreases
__global__ void count_decreasing_kernel(const int64_t *rowptr, long M, int *flag)
{
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < M && rowptr[i + 1] < rowptr[i]) atomicAdd(flag, 1);
}

// Checks a CSR (int64 rowptr, int32 columns) on the device: rowptr starts at
// 0 and never decreases, every column is in [0, ncols), nnz < 2^31.  Returns nnz.
long validate_csr(rbffd_context *c, long M, const int64_t *rowptr, const int *col, long ncols, const char *what)
{
    hipStream_t st = c->stream;
    int64_t ends[2] = {0, 0};
    RBFFD_HIP(hipMemcpyAsync(&ends[0], rowptr, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    RBFFD_HIP(hipMemcpyAsync(&ends[1], rowptr + M, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    RBFFD_HIP(hipStreamSynchronize(st));
    const long nnz = (long)ends[1];
    if (ends[0] != 0 || nnz < 0) throw Error{RBFFD_ERR_ARG, std::string(what) + ": rowptr must start at 0 and end at nnz"};
    if (nnz > INT_MAX) throw Error{RBFFD_ERR_ARG, std::string(what) + ": more than 2^31 - 1 entries"};
    if (nnz == 0) return 0;
    int *flag = (int *)ws(c, S_FLAG, sizeof(int) * 8);
    RBFFD_HIP(hipMemsetAsync(flag, 0, sizeof(int) * 3, st));
    max_kernel<<<(unsigned)std::min<long>(2048, nblk(nnz)), 256, 0, st>>>(col, nnz, flag);
    count_negative_kernel<<<nblk(nnz), 256, 0, st>>>(col, nnz, flag + 1);
    count_decreasing_kernel<<<nblk(M), 256, 0, st>>>(rowptr, M, flag + 2);
    int hf[3] = {0, 0, 0};
    RBFFD_HIP(hipMemcpyAsync(hf, flag, sizeof(hf), hipMemcpyDeviceToHost, st));
    RBFFD_HIP(hipStreamSynchronize(st));
    if (hf[1] > 0 || hf[0] >= ncols) throw Error{RBFFD_ERR_ARG, std::string(what) + ": column index out of range"};
    if (hf[2] > 0) throw Error{RBFFD_ERR_ARG, std::string(what) + ": rowptr decreases"};
    return nnz;
}

// Indices all in [0, n); flag: 2 device ints of scratch (NULL: the context's)
void validate_indices(rbffd_context *c, const int *idx, long count, long n, const char *what, int *flag = nullptr)
{
    if (count <= 0) return;
    hipStream_t st = c->stream;
    if (!flag) flag = (int *)ws(c, S_FLAG, sizeof(int) * 8);
    rbffd::launch_fill_i32(flag, 2, 0, st);
    max_kernel<<<(unsigned)std::min<long>(2048, nblk(count)), 256, 0, st>>>(idx, count, flag);
    count_negative_kernel<<<nblk(count), 256, 0, st>>>(idx, count, flag + 1);
    int hf[2] = {0, 0};
    RBFFD_HIP(hipMemcpyAsync(hf, flag, sizeof(hf), hipMemcpyDeviceToHost, st));
    RBFFD_HIP(hipStreamSynchronize(st));
    if (hf[1] > 0 || hf[0] >= n) throw Error{RBFFD_ERR_ARG, std::string(what) + ": index out of range"};
}

// Iterations per recorded graph.  Even: the two gamma slots alternate by
// iteration parity, and every replay starts at an even iteration.
constexpr long kLsqChunk = 32;

struct LsqState {
    long M, N, nnz;
    const int64_t *rowptr;
    const int *col;
    const double *dval;  // D C
    const int64_t *tptr;
    const int *tcol;
    const double *tval;  // (D C)^T
    double *scale, *z, *p, *s, *r, *q;
    double *g;           // g[0], g[1]: gamma by iteration parity; g[2]: |q|^2
    double *partial;
    unsigned *count;
};

// CGLS iteration k (its gamma is g[k & 1]; the next one goes to g[(k + 1) & 1])
void lsq_iteration(const LsqState &S, long k, hipStream_t st)
{
    const double *gam = S.g + (k & 1);
    double *gam_new = S.g + ((k + 1) & 1);
    rbffd::launch_lsq_spmv(S.M, S.nnz, S.rowptr, S.col, S.dval, S.p, S.q, {S.partial, S.count, S.g + 2}, st);
    rbffd::launch_lsq_update(S.N, S.M, gam, S.g + 2, S.p, S.q, S.z, S.r, st);
    rbffd::launch_lsq_spmv(S.N, S.nnz, S.tptr, S.tcol, S.tval, S.r, S.s, {S.partial, S.count, gam_new}, st);
    rbffd::launch_lsq_direction(S.N, gam, gam_new, S.s, S.p, st);
}

// Largest dense copy (entries) the automatic choice sends to the direct method.
constexpr long kDenseAutoEntries = 1L << 28;

void lsq_impl(rbffd_context *c, long M, long N, const int64_t *rowptr, const int *col, const double *vals,
              const double *b, double *x, int method, double tol, long max_iters, int64_t *iters_out,
              double *relres_out)
{
    hipStream_t st = c->stream;
    if (iters_out) *iters_out = 0;
    if (relres_out) *relres_out = 0.0;
    if (M < N) throw Error{RBFFD_ERR_ARG, "lsq: D needs at least as many rows as columns"};
    if (M > INT_MAX || N > INT_MAX) throw Error{RBFFD_ERR_ARG, "lsq: more than 2^31 - 1 rows"};
    if (!(tol >= 0.0) || max_iters < 0) throw Error{RBFFD_ERR_ARG, "lsq: tol and max_iters must be >= 0"};
    if (N == 0) return;
    const long nnz = validate_csr(c, M, rowptr, col, N, "lsq");
    if (nnz == 0) return;  // D = 0: every x is a least-squares solution; x stays the start

    if (method == RBFFD_SOLVE_AUTO)
        method = (rbffd::dense_solver_available() && M * N <= kDenseAutoEntries) ? RBFFD_SOLVE_QR : RBFFD_SOLVE_CGLS;
    if (method == RBFFD_SOLVE_QR) {
        if (M * N > (long)INT_MAX) throw Error{RBFFD_ERR_ARG, "lsq: the direct method needs M N < 2^31"};
        double *A = (double *)ws(c, S_LSQ_DENSE, sizeof(double) * ((size_t)M * N + M));
        int *info = (int *)ws(c, S_LSQ_INFO, sizeof(int) * 4);
        const int r = rbffd::dense_lsq(&c->solver, M, N, rowptr, col, vals, b, x, A, A + (size_t)M * N, info, st);
        if (r > 0) throw Error{RBFFD_ERR_SINGULAR, "lsq: D is rank deficient (zero diagonal " + std::to_string(r) + " in R)"};
        return;
    }

    // the transpose as its own CSR, rows ascending inside each transposed row
    int *row_of = (int *)ws(c, S_LSQ_ROWOF, sizeof(int) * nnz);
    int *counts = (int *)ws(c, S_CSC_COUNTS, sizeof(int) * (N + 1));
    int *ptr = (int *)ws(c, S_CSC_PTR, sizeof(int) * (N + 1));
    int *fill = (int *)ws(c, S_CSC_FILL, sizeof(int) * N);
    int *perm_in = (int *)ws(c, S_CSC_PERM_IN, sizeof(int) * nnz);
    int *perm = (int *)ws(c, S_CSC_PERM, sizeof(int) * nnz);
    int *entry = (int *)ws(c, S_CSC_ROW, sizeof(int) * nnz);
    int *tcol = (int *)ws(c, S_LSQ_TCOL, sizeof(int) * nnz);
    int64_t *tptr = (int64_t *)ws(c, S_LSQ_TPTR, sizeof(int64_t) * (N + 1));
    const size_t nf = 2 * (size_t)nnz + 4 * (size_t)N + 2 * (size_t)M + rbffd::kLsqBlocks + 4;
    double *f = (double *)ws(c, S_LSQ_F64, sizeof(double) * nf);
    unsigned *count = (unsigned *)ws(c, S_LSQ_COUNT, sizeof(unsigned) * 4);
    LsqState S;
    S.M = M; S.N = N; S.nnz = nnz; S.rowptr = rowptr; S.col = col;
    double *dval = f, *tval = dval + nnz;
    S.scale = tval + nnz; S.z = S.scale + N; S.p = S.z + N; S.s = S.p + N;
    S.r = S.s + N; S.q = S.r + M; S.partial = S.q + M; S.g = S.partial + rbffd::kLsqBlocks;
    S.dval = dval; S.tptr = tptr; S.tcol = tcol; S.tval = tval; S.count = count;

    rbffd::launch_row_of(rowptr, M, row_of, st);
    rbffd::launch_fill_i32(counts, N + 1, 0, st);
    rbffd::launch_fill_i32(fill, N, 0, st);
    rbffd::launch_fill_i32((int *)count, 4, 0, st);
    rbffd::launch_csc_count(col, nnz, counts, st);
    const size_t tb = rbffd::scan_tmp_bytes(N + 1);
    void *tmp = ws(c, S_SCAN_TMP2, tb);
    rbffd::launch_exclusive_scan(counts, ptr, N + 1, tmp, tb, st);
    rbffd::launch_csc_scatter(col, nnz, 1, ptr, fill, perm_in, st);
    rbffd::launch_csc_rank(col, nnz, 1, ptr, perm_in, perm, entry, st);
    rbffd::launch_gather_i32(row_of, perm, nnz, tcol, st);
    rbffd::launch_i32_to_i64(ptr, N + 1, tptr, 0, st);
    rbffd::launch_gather_f64(vals, perm, nnz, tval, st);       // D^T
    rbffd::launch_col_scale(tptr, N, tval, S.scale, st);
    rbffd::launch_scale_cols(col, nnz, S.scale, vals, dval, st);
    rbffd::launch_gather_f64(dval, perm, nnz, tval, st);       // (D C)^T

    // r = b - D x0, z = 0, s = (D C)^T r, gamma0 = |s|^2, p = s
    rbffd::launch_copy_f64(b, M, S.r, st);
    rbffd::launch_spmv(M, rowptr, col, vals, 0, 1, x, S.r, -1.0, 1.0, 0, nullptr, st);
    rbffd::launch_fill_f64(S.z, N, 0.0, st);
    rbffd::launch_fill_f64(S.g, 4, 0.0, st);
    rbffd::launch_lsq_spmv(N, nnz, tptr, tcol, tval, S.r, S.s, {S.partial, count, S.g}, st);
    rbffd::launch_copy_f64(S.s, N, S.p, st);
    RBFFD_HIP(hipGetLastError());
    double g0 = 0.0;
    RBFFD_HIP(hipMemcpyAsync(&g0, S.g, sizeof(double), hipMemcpyDeviceToHost, st));
    RBFFD_HIP(hipStreamSynchronize(st));
    if (!std::isfinite(g0)) throw Error{RBFFD_ERR_NONFINITE, "lsq: non-finite values in D, b or x"};

    long done = 0;
    double rel = 0.0;
    if (g0 > 0.0 && max_iters > 0) {
        rel = 1.0;
        // a chunk of iterations recorded once and replayed (the stream must
        // be capturable: not the null stream)
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        if (st && max_iters >= kLsqChunk) {
            RBFFD_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                for (long k = 0; k < kLsqChunk; ++k) lsq_iteration(S, k, st);
                RBFFD_HIP(hipGetLastError());
            } catch (...) {
                // leave the stream out of capture mode
                (void)hipStreamEndCapture(st, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            RBFFD_HIP(hipStreamEndCapture(st, &graph));
            const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(graph);
                throw Error{RBFFD_ERR_RUNTIME, std::string("lsq: hipGraphInstantiate: ") + hipGetErrorString(e)};
            }
        }
        try {
            while (done < max_iters && rel > tol) {
                const long k = std::min(kLsqChunk, max_iters - done);
                if (exec && k == kLsqChunk)
                    RBFFD_HIP(hipGraphLaunch(exec, st));
                else
                    for (long j = 0; j < k; ++j) lsq_iteration(S, done + j, st);
                done += k;
                double g = 0.0;
                RBFFD_HIP(hipMemcpyAsync(&g, S.g + (done & 1), sizeof(double), hipMemcpyDeviceToHost, st));
                RBFFD_HIP(hipStreamSynchronize(st));
                if (!std::isfinite(g)) throw Error{RBFFD_ERR_NONFINITE, "lsq: non-finite iterate"};
                rel = std::sqrt(g / g0);
            }
        } catch (...) {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        if (exec) RBFFD_HIP(hipGraphExecDestroy(exec));
        if (graph) RBFFD_HIP(hipGraphDestroy(graph));
    }
    rbffd::launch_lsq_finish(N, S.scale, S.z, x, st);
    RBFFD_HIP(hipGetLastError());
    RBFFD_HIP(hipStreamSynchronize(st));
    if (iters_out) *iters_out = done;
    if (relres_out) *relres_out = rel;
}

// ------------------------------------------------------------ batch
// Per set: 8 flag words.  [0] bit 0: X not finite; [1] kNN result slots at -1;
// [2 + s] rows with status s (1 singular, 2 non-finite weights).
constexpr int kSetFlags = 8;

void batch_enqueue(rbffd_context *c, const rbffd_batch_set *sets, int nsets, const std::vector<int> &lane_of,
                   int nops, const int *ops, const int *opk, int *flags, bool capture)
{
    hipStream_t main = c->stream;
    rbffd::launch_fill_i32(flags, (long)kSetFlags * nsets, 0, main);
    if (capture) {
        RBFFD_HIP(hipEventRecord(c->fork_ev, main));
        for (int l = 0; l < kLanes; ++l) RBFFD_HIP(hipStreamWaitEvent(c->lane_stream[l], c->fork_ev, 0));
    }
    for (int i = 0; i < nsets; ++i) {
        const rbffd_batch_set &S = sets[i];
        const int l = lane_of[i];
        hipStream_t s = capture ? c->lane_stream[l] : main;
        rbffd_arena &A = c->lane_arena[l];
        int *st = (int *)grow(A.b[A_STENCILS], sizeof(int) * (size_t)S.N * S.n, !capture);
        int *status = S.status ? S.status : (int *)grow(A.b[A_STATUS], sizeof(int) * (size_t)S.N, !capture);
        int *f = flags + (long)kSetFlags * i;
        knn_enqueue(A, !capture, S.X, S.N, S.X, S.N, S.n, S.xgroup, S.qgroup, st, nullptr, f, s);
        Workspace w{c};
        w.may_alloc = !capture;
        solves_enqueue(c, w, S.N, S.X, S.X, S.N, st, S.n, nullptr, nullptr, 0, S.p, S.polydeg, 1, nops, ops, opk,
                       RBFFD_PREC_FP64, S.colidx, S.vals, status, s, !capture, A.b[A_GEN_SCRATCH]);
        status_count_kernel<<<nblk(S.N), 256, 0, s>>>(status, S.N, f + 2);
        RBFFD_HIP(hipGetLastError());
    }
    if (capture) {
        for (int l = 0; l < kLanes; ++l) {
            RBFFD_HIP(hipEventRecord(c->join_ev[l], c->lane_stream[l]));
            RBFFD_HIP(hipStreamWaitEvent(main, c->join_ev[l], 0));
        }
    }
}

// Every context-owned buffer a captured batch can reference.  A buffer that
// grows is freed and reallocated, so a graph recorded before that would
// replay into freed memory: batch_impl replays only while this list is
// unchanged.
std::vector<const void *> batch_buffers(const rbffd_context *c)
{
    std::vector<const void *> v;
    for (const auto &A : c->lane_arena)
        for (const auto &b : A.b) v.push_back(b.p);
    for (const auto &b : c->bufs) v.push_back(b.p);
    return v;
}

std::vector<int64_t> batch_key(const rbffd_batch_set *sets, int nsets, int nops, const int *ops, const int *opk)
{
    std::vector<int64_t> k{nsets, nops};
    for (int o = 0; o < nops; ++o) {
        k.push_back(ops[o]);
        k.push_back(opk ? opk[o] : 0);
    }
    for (int i = 0; i < nsets; ++i) {
        const auto &S = sets[i];
        for (int64_t v : {(int64_t)(intptr_t)S.X, S.N, (int64_t)S.p, (int64_t)S.n, (int64_t)S.polydeg,
                          (int64_t)(intptr_t)S.xgroup, (int64_t)(intptr_t)S.qgroup,
                          (int64_t)(intptr_t)S.colidx, (int64_t)(intptr_t)S.status})
            k.push_back(v);
        for (int o = 0; o < nops; ++o) k.push_back((int64_t)(intptr_t)S.vals[o]);
    }
    return k;
}

void batch_impl(rbffd_context *c, int nsets, const rbffd_batch_set *sets, int nops, const int *ops,
                const int *opk, int use_graph, int32_t *set_status, rbffd_stats *stats)
{
    check_ctx(c);
    if (nsets < 1 || !sets) throw Error{RBFFD_ERR_ARG, "batch: no sets"};
    check_ops(nops, ops, opk);
    std::vector<double> work(nsets);
    for (int i = 0; i < nsets; ++i) {
        const auto &S = sets[i];
        if (!S.X || !S.colidx) throw Error{RBFFD_ERR_ARG, "batch: null X or colidx in set " + std::to_string(i)};
        for (int o = 0; o < nops; ++o)
            if (!S.vals[o]) throw Error{RBFFD_ERR_ARG, "batch: null vals in set " + std::to_string(i)};
        if (S.n < 1 || S.N < S.n) throw Error{RBFFD_ERR_ARG, "batch: set " + std::to_string(i) + " has fewer nodes than n"};
        if ((S.xgroup == nullptr) != (S.qgroup == nullptr))
            throw Error{RBFFD_ERR_ARG, "batch: xgroup and qgroup go together"};
        check_weights_args(S.n, S.p, S.polydeg, 1, nops, ops, opk, RBFFD_PREC_FP64);
        const double ns = S.n + n_monomials(S.polydeg);
        work[i] = (double)S.N * ns * ns * ns;
    }
    // longest first onto the least loaded branch
    std::vector<int> order(nsets), lane_of(nsets);
    for (int i = 0; i < nsets; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
    double load[kLanes] = {};
    for (int i : order) {
        int l = (int)(std::min_element(load, load + kLanes) - load);
        lane_of[i] = l;
        load[l] += work[i];
    }
    int *flags = (int *)ws(c, S_BATCH_FLAGS, sizeof(int) * kSetFlags * nsets);
    for (int e = 0; e < 2; ++e)
        if (!c->ev[e]) RBFFD_HIP(hipEventCreate(&c->ev[e]));
    hipStream_t main = c->stream;
    auto key = batch_key(sets, nsets, nops, ops, opk);
    RBFFD_HIP(hipEventRecord(c->ev[0], main));
    if (use_graph && c->batch_exec && key == c->batch_key && batch_buffers(c) == c->batch_bufs) {
        RBFFD_HIP(hipGraphLaunch(c->batch_exec, main));
    } else {
        batch_enqueue(c, sets, nsets, lane_of, nops, ops, opk, flags, false);  // also tunes the layouts
        if (use_graph) {
            if (c->batch_exec) { RBFFD_HIP(hipGraphExecDestroy(c->batch_exec)); c->batch_exec = nullptr; }
            if (c->batch_graph) { RBFFD_HIP(hipGraphDestroy(c->batch_graph)); c->batch_graph = nullptr; }
            c->batch_key.clear();
            c->batch_bufs.clear();
            for (int l = 0; l < kLanes; ++l) {
                if (!c->lane_stream[l]) RBFFD_HIP(hipStreamCreateWithFlags(&c->lane_stream[l], hipStreamNonBlocking));
                if (!c->join_ev[l]) RBFFD_HIP(hipEventCreateWithFlags(&c->join_ev[l], hipEventDisableTiming));
            }
            if (!c->fork_ev) RBFFD_HIP(hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
            // the eager pass above must finish before capture: the graph reuses its buffers
            RBFFD_HIP(hipStreamSynchronize(main));
            RBFFD_HIP(hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal));
            hipGraph_t g = nullptr;
            try {
                batch_enqueue(c, sets, nsets, lane_of, nops, ops, opk, flags, true);
            } catch (...) {
                (void)hipStreamEndCapture(main, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            RBFFD_HIP(hipStreamEndCapture(main, &g));
            c->batch_graph = g;
            RBFFD_HIP(hipGraphInstantiate(&c->batch_exec, g, nullptr, nullptr, 0));
            c->batch_key = key;
            c->batch_bufs = batch_buffers(c);
        }
    }
    RBFFD_HIP(hipEventRecord(c->ev[1], main));
    std::vector<int> hf((size_t)kSetFlags * nsets);
    RBFFD_HIP(hipMemcpyAsync(hf.data(), flags, sizeof(int) * hf.size(), hipMemcpyDeviceToHost, main));
    RBFFD_HIP(hipStreamSynchronize(main));
    long nsing = 0, nnonf = 0, rows = 0;
    int first_code = RBFFD_OK, nbad = 0;
    for (int i = 0; i < nsets; ++i) {
        const int *f = &hf[(size_t)kSetFlags * i];
        int st = 0, code = RBFFD_OK;
        if (f[0] & 1) { st = 2; code = RBFFD_ERR_NONFINITE; }
        else if (f[1] != 0) { st = 3; code = RBFFD_ERR_ARG; }
        else if (f[3] != 0) { st = 1; code = RBFFD_ERR_SINGULAR; }
        else if (f[4] != 0) { st = 2; code = RBFFD_ERR_NONFINITE; }
        if (set_status) set_status[i] = st;
        nsing += f[3];
        nnonf += f[4];
        rows += sets[i].N;
        if (code != RBFFD_OK) {
            ++nbad;
            if (first_code == RBFFD_OK) first_code = code;
        }
    }
    if (stats) {
        *stats = rbffd_stats{};
        float ms = 0;
        RBFFD_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        stats->rows = rows;
        stats->singular = nsing;
        stats->nonfinite = nnonf;
        stats->ms_solve = ms;
    }
    if (first_code != RBFFD_OK)
        throw Error{first_code, std::to_string(nbad) + " of " + std::to_string(nsets) +
                                    " sets failed (see set_status)"};
}

}  // namespace

// =================================================================== C ABI
extern "C" {

int rbffd_version(void) { return RBFFD_VERSION; }

const char *rbffd_last_error(void) { return g_last_error.c_str(); }

int rbffd_create(int device, rbffd_context **out)
{
    return guarded([&] {
        if (!out) throw Error{RBFFD_ERR_ARG, "null output"};
        int count = 0;
        RBFFD_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw Error{RBFFD_ERR_ARG, "no such device"};
        RBFFD_HIP(hipSetDevice(device));
        // the no-allocation entry point's array, looked up with the device
        // current; without it that entry point reports an error and the rest
        // of the library works as before
        char *sp = nullptr;
        try {
            sp = scratch_address(device);
        } catch (const Error &) {
            sp = nullptr;
        }
        auto *c = new rbffd_context();
        c->device = device;
        c->noalloc = sp;
        const hipError_t e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            throw Error{RBFFD_ERR_RUNTIME, std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e)};
        }
        c->stream = c->own_stream;
        *out = c;
    });
}

int rbffd_destroy(rbffd_context *c)
{
    return guarded([&] {
        if (!c) return;
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        for (auto &st : c->lane_stream)
            if (st) (void)hipStreamSynchronize(st);
        if (c->batch_exec) (void)hipGraphExecDestroy(c->batch_exec);
        if (c->batch_graph) (void)hipGraphDestroy(c->batch_graph);
        rbffd::dense_solver_destroy(c->solver);
        for (auto &b : c->bufs)
            if (b.p) (void)hipFree(b.p);
        for (auto &b : c->knn_arena.b)
            if (b.p) (void)hipFree(b.p);
        for (auto &A : c->lane_arena)
            for (auto &b : A.b)
                if (b.p) (void)hipFree(b.p);
        for (auto &e : c->ev)
            if (e) (void)hipEventDestroy(e);
        for (auto &e : c->tune_ev)
            if (e) (void)hipEventDestroy(e);
        c->pool.reset();
        for (auto &b : c->stage)
            if (b) (void)hipHostFree(b);
        for (auto &e : c->stage_ev)
            if (e) (void)hipEventDestroy(e);
        if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
        for (auto &e : c->join_ev)
            if (e) (void)hipEventDestroy(e);
        for (auto &st : c->lane_stream)
            if (st) (void)hipStreamDestroy(st);
        if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
        delete c;
    });
}

int rbffd_set_stream(rbffd_context *c, void *stream)
{
    return guarded([&] {
        check_ctx(c);
        c->stream = stream ? (hipStream_t)stream : c->own_stream;
    });
}

int rbffd_synchronize(rbffd_context *c)
{
    return guarded([&] {
        check_ctx(c);
        RBFFD_HIP(hipStreamSynchronize(c->stream));
    });
}

int rbffd_device_malloc(rbffd_context *c, int64_t bytes, void **ptr)
{
    return guarded([&] {
        check_ctx(c);
        if (!ptr || bytes < 0) throw Error{RBFFD_ERR_ARG, "bad allocation request"};
        *ptr = nullptr;
        RBFFD_HIP(hipMalloc(ptr, (size_t)(bytes > 0 ? bytes : 16)));
    });
}

int rbffd_device_free(rbffd_context *c, void *ptr)
{
    return guarded([&] {
        check_ctx(c);
        if (ptr) {
            RBFFD_HIP(hipStreamSynchronize(c->stream));
            RBFFD_HIP(hipFree(ptr));
        }
    });
}

static void copy_sync(rbffd_context *c, void *dst, const void *src, int64_t bytes, hipMemcpyKind kind)
{
    check_ctx(c);
    if (bytes < 0 || ((!dst || !src) && bytes > 0)) throw Error{RBFFD_ERR_ARG, "bad copy"};
    if (bytes == 0) return;
    RBFFD_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, kind, c->stream));
    RBFFD_HIP(hipStreamSynchronize(c->stream));
}

int rbffd_memcpy_h2d(rbffd_context *c, void *dst, const void *src, int64_t bytes)
{
    return guarded([&] { copy_sync(c, dst, src, bytes, hipMemcpyHostToDevice); });
}

int rbffd_memcpy_d2h(rbffd_context *c, void *dst, const void *src, int64_t bytes)
{
    return guarded([&] { copy_sync(c, dst, src, bytes, hipMemcpyDeviceToHost); });
}

int rbffd_memcpy_d2d(rbffd_context *c, void *dst, const void *src, int64_t bytes)
{
    return guarded([&] { copy_sync(c, dst, src, bytes, hipMemcpyDeviceToDevice); });
}

int rbffd_memset_device(rbffd_context *c, void *dst, int32_t value, int64_t bytes)
{
    return guarded([&] {
        check_ctx(c);
        if (!dst || bytes < 0) throw Error{RBFFD_ERR_ARG, "bad memset"};
        RBFFD_HIP(hipMemsetAsync(dst, value, (size_t)bytes, c->stream));
        RBFFD_HIP(hipStreamSynchronize(c->stream));
    });
}

int rbffd_timer_start(rbffd_context *c)
{
    return guarded([&] {
        check_ctx(c);
        for (int e = 6; e < 8; ++e)
            if (!c->ev[e]) RBFFD_HIP(hipEventCreate(&c->ev[e]));
        RBFFD_HIP(hipEventRecord(c->ev[6], c->stream));
    });
}

int rbffd_timer_stop(rbffd_context *c, float *ms)
{
    return guarded([&] {
        check_ctx(c);
        if (!ms || !c->ev[6]) throw Error{RBFFD_ERR_ARG, "timer not started"};
        RBFFD_HIP(hipEventRecord(c->ev[7], c->stream));
        RBFFD_HIP(hipEventSynchronize(c->ev[7]));
        RBFFD_HIP(hipEventElapsedTime(ms, c->ev[6], c->ev[7]));
    });
}

int rbffd_device_info(rbffd_context *c, char *name, int32_t name_len, int32_t *cus, char *arch, int32_t arch_len)
{
    return guarded([&] {
        check_ctx(c);
        hipDeviceProp_t p;
        RBFFD_HIP(hipGetDeviceProperties(&p, c->device));
        if (name && name_len > 0) { strncpy(name, p.name, name_len - 1); name[name_len - 1] = 0; }
        if (cus) *cus = p.multiProcessorCount;
        if (arch && arch_len > 0) { strncpy(arch, p.gcnArchName, arch_len - 1); arch[arch_len - 1] = 0; }
    });
}

int rbffd_knn_device(rbffd_context *c, const double *X, int64_t N, const double *Q, int64_t M, int32_t k,
                     const int32_t *xgroup, const int32_t *qgroup, int32_t *idx, double *d2)
{
    return guarded([&] {
        check_ctx(c);
        if (!X || !Q || !idx) throw Error{RBFFD_ERR_ARG, "null pointer"};
        knn_impl(c, X, N, Q, M, k, xgroup, qgroup, idx, d2);
    });
}

int rbffd_weights_device(rbffd_context *c, const double *X, int64_t N, const double *Y, int64_t M,
                         const int32_t *stencils, int32_t n, const int32_t *y2x, int32_t p, int32_t polydeg,
                         int32_t method, int32_t nops, const int32_t *ops, const int32_t *opk,
                         int32_t precision, int32_t *colidx, double *const *vals, int32_t *status)
{
    return guarded([&] {
        check_ctx(c);
        if (!X || !Y || !stencils || !ops || !vals) throw Error{RBFFD_ERR_ARG, "null pointer"};
        if (N < 1 || M < 0 || N > INT_MAX || M > INT_MAX) throw Error{RBFFD_ERR_ARG, "weights: bad sizes"};
        for (int o = 0; o < nops && o < RBFFD_MAX_OPS; ++o)
            if (!vals[o]) throw Error{RBFFD_ERR_ARG, "null value array"};
        if (M == 0) return;
        if (n >= 1 && (long)M * n > INT_MAX) throw Error{RBFFD_ERR_ARG, "weights: more than 2^31 - 1 entries"};
        // every index the kernels follow stays inside X
        const long nst = (y2x ? N : M) * (long)n;
        if (n >= 1) validate_indices(c, stencils, nst, N, "weights stencils");
        if (y2x) validate_indices(c, y2x, M, N, "weights y2x");
        weights_impl(c, X, N, Y, M, stencils, n, y2x, nullptr, 0, p, polydeg, method, nops, ops, opk, precision,
                     colidx, vals, status);
    });
}

int rbffd_weights_device_noalloc(rbffd_context *c, const double *X, int64_t N, const double *Y, int64_t M,
                                 const int32_t *stencils, int32_t n, const int32_t *y2x, int32_t p,
                                 int32_t polydeg, int32_t method, int32_t nops, const int32_t *ops,
                                 const int32_t *opk, int32_t precision, int32_t *colidx, double *const *vals,
                                 int32_t *status)
{
    return guarded([&] {
        check_ctx(c);
        if (!X || !Y || !stencils || !ops || !vals) throw Error{RBFFD_ERR_ARG, "null pointer"};
        if (N < 1 || M < 0 || N > INT_MAX || M > INT_MAX) throw Error{RBFFD_ERR_ARG, "weights: bad sizes"};
        for (int o = 0; o < nops && o < RBFFD_MAX_OPS; ++o)
            if (!vals[o]) throw Error{RBFFD_ERR_ARG, "null value array"};
        if (!c->noalloc) throw Error{RBFFD_ERR_RUNTIME, "no static workspace on this device"};
        if (M == 0) return;
        if (n >= 1 && (long)M * n > INT_MAX) throw Error{RBFFD_ERR_ARG, "weights: more than 2^31 - 1 entries"};
        // every buffer below comes from the static array; nothing is allocated
        Workspace w{c, c->noalloc, kNoAllocBytes, 0, false};
        int *flag = (int *)w.take(sizeof(int) * 8);
        const long nst = (y2x ? N : M) * (long)n;
        if (n >= 1) validate_indices(c, stencils, nst, N, "weights stencils", flag);
        if (y2x) validate_indices(c, y2x, M, N, "weights y2x", flag);
        weights_impl(c, X, N, Y, M, stencils, n, y2x, nullptr, 0, p, polydeg, method, nops, ops, opk, precision,
                     colidx, vals, status, nullptr, &w);
    });
}

int rbffd_generate_device(rbffd_context *c, const double *X, int64_t N, const double *Y, int64_t M, int32_t p,
                          int32_t n, int32_t polydeg, const rbffd_boundary *bnd, int32_t nops,
                          const int32_t *ops, const int32_t *opk, int32_t precision, int32_t *colidx,
                          double *const *vals, int32_t *stencils_out, int32_t *y2x_out, rbffd_stats *stats)
{
    return guarded([&] {
        check_ctx(c);
        if (!X || !ops || !vals) throw Error{RBFFD_ERR_ARG, "null pointer"};
        check_ops(nops, ops, opk);
        for (int o = 0; o < nops; ++o)
            if (!vals[o]) throw Error{RBFFD_ERR_ARG, "null value array"};
        generate_impl(c, X, N, Y, M, p, n, polydeg, bnd, nops, ops, opk, precision, colidx, vals, stencils_out,
                      y2x_out, stats, nullptr);
    });
}

int rbffd_csr_to_csc_device(rbffd_context *c, int64_t M, int32_t n, int64_t ncols, const int32_t *colidx,
                            int32_t nops, const double *const *csr_vals, int64_t *colptr, int64_t *rowval,
                            double *const *csc_vals, int32_t index_base, int64_t *ncols_used)
{
    return guarded([&] {
        check_ctx(c);
        if (!colidx || !colptr || !rowval || (nops > 0 && (!csr_vals || !csc_vals)))
            throw Error{RBFFD_ERR_ARG, "null pointer"};
        if (M < 1 || n < 1) throw Error{RBFFD_ERR_ARG, "empty matrix"};
        csc_impl(c, M, n, ncols, colidx, nops, csr_vals, colptr, rowval, csc_vals, index_base, ncols_used);
    });
}

int rbffd_lsq_solve_device(rbffd_context *c, int64_t M, int64_t N, const int64_t *rowptr, const int32_t *col,
                           const double *vals, const double *b, double *x, int32_t method, double tol,
                           int64_t max_iters, int64_t *iters, double *relres)
{
    return guarded([&] {
        check_ctx(c);
        if (M < 0 || N < 0) throw Error{RBFFD_ERR_ARG, "negative size"};
        if (N > 0 && (!rowptr || !col || !vals || !b || !x)) throw Error{RBFFD_ERR_ARG, "null pointer"};
        if (method < RBFFD_SOLVE_AUTO || method > RBFFD_SOLVE_QR) throw Error{RBFFD_ERR_ARG, "unknown solve method"};
        lsq_impl(c, M, N, rowptr, col, vals, b, x, method, tol, max_iters, iters, relres);
    });
}

int rbffd_spmv_device(rbffd_context *c, int64_t M, const int64_t *rowptr, const int32_t *col, const void *vals,
                      int32_t vals_type, int32_t F, const double *x, double *y, double alpha, double beta,
                      int32_t uniform_n, const int32_t *row_order)
{
    return guarded([&] {
        check_ctx(c);
        if (!col || !vals || !x || !y) throw Error{RBFFD_ERR_ARG, "null pointer"};
        if (uniform_n <= 0 && !rowptr) throw Error{RBFFD_ERR_ARG, "rowptr needed for non-uniform rows"};
        if (vals_type != RBFFD_F64 && vals_type != RBFFD_F32) throw Error{RBFFD_ERR_ARG, "unknown value type"};
        rbffd::launch_spmv(M, rowptr, col, vals, vals_type == RBFFD_F32, F, x, y, alpha, beta, uniform_n,
                           row_order, c->stream);
        RBFFD_HIP(hipGetLastError());
    });
}

struct rbffd_spmv_plan {
    int device = 0;
    rbffd::SpmvPlan p;
};

int rbffd_spmv_plan_create(rbffd_context *c, int64_t M, int64_t N, const int64_t *rowptr, const int32_t *col,
                           const void *vals, int32_t vals_type, int32_t uniform_n, const int32_t *row_order,
                           rbffd_spmv_plan **plan, int64_t *info)
{
    return guarded([&] {
        check_ctx(c);
        if (!plan) throw Error{RBFFD_ERR_ARG, "spmv plan: null pointer"};
        *plan = nullptr;
        if (M < 0 || N < 0 || (M > 0 && (!col || !vals))) throw Error{RBFFD_ERR_ARG, "spmv plan: bad arguments"};
        if (uniform_n <= 0 && !rowptr) throw Error{RBFFD_ERR_ARG, "spmv plan: rowptr or uniform_n needed"};
        if (vals_type != RBFFD_F64 && vals_type != RBFFD_F32) throw Error{RBFFD_ERR_ARG, "spmv plan: value type"};
        if (M >= (int64_t)INT_MAX) throw Error{RBFFD_ERR_ARG, "spmv plan: needs M < 2^31 - 1"};
        auto q = std::make_unique<rbffd_spmv_plan>();
        q->device = c->device;
        rbffd::spmv_plan_build(q->p, M, N, rowptr, col, vals, vals_type == RBFFD_F32, uniform_n, row_order,
                               c->stream);
        if (info) {
            info[0] = q->p.nblocks;
            info[1] = q->p.hmax;
            info[2] = q->p.overflow_blocks;
            info[3] = q->p.nnz;
        }
        *plan = q.release();
    });
}

int rbffd_spmv_plan_apply(rbffd_context *c, const rbffd_spmv_plan *plan, int32_t F, const double *x, double *y,
                          double alpha, double beta)
{
    return guarded([&] {
        check_ctx(c);
        if (!plan || (plan->p.M > 0 && (!x || !y))) throw Error{RBFFD_ERR_ARG, "spmv plan: null pointer"};
        if (plan->device != c->device) throw Error{RBFFD_ERR_ARG, "spmv plan: made on another device"};
        if (F != 1 && F != 2 && F != 4 && F != 8) throw Error{RBFFD_ERR_UNSUPPORTED, "spmv plan: fields must be 1, 2, 4 or 8"};
        rbffd::launch_spmv_plan(plan->p, F, x, y, alpha, beta, c->stream);
        RBFFD_HIP(hipGetLastError());
    });
}

int rbffd_spmv_plan_destroy(rbffd_spmv_plan *plan)
{
    return guarded([&] {
        if (!plan) return;
        RBFFD_HIP(hipSetDevice(plan->device));
        rbffd::spmv_plan_free(plan->p);
        delete plan;
    });
}

int rbffd_ssprk3_plan_device(rbffd_context *c, const rbffd_spmv_plan *plan, int32_t F, double *u, double *work,
                             const uint8_t *fixed, double dt, int64_t steps)
{
    return guarded([&] {
        check_ctx(c);
        if (!plan || steps < 0 || (plan->p.M > 0 && (!u || !work))) throw Error{RBFFD_ERR_ARG, "ssprk3: bad arguments"};
        if (plan->device != c->device) throw Error{RBFFD_ERR_ARG, "ssprk3: plan made on another device"};
        if (plan->p.M != plan->p.N) throw Error{RBFFD_ERR_ARG, "ssprk3: the operator must be square"};
        if (F != 1 && F != 2 && F != 4 && F != 8) throw Error{RBFFD_ERR_UNSUPPORTED, "ssprk3: fields must be 1, 2, 4 or 8"};
        hipStream_t s = c->stream;
        const rbffd::SpmvPlan &P = plan->p;
        double *u1 = work, *u2 = work + (size_t)P.M * F;
        for (int64_t k = 0; k < steps; ++k) {
            // the stages of rbffd_ssprk3_device
            rbffd::launch_rk_stage_plan(P, F, u, u, u1, fixed, 0.0, 1.0, dt, s);
            rbffd::launch_rk_stage_plan(P, F, u1, u, u2, fixed, 0.75, 0.25, 0.25 * dt, s);
            rbffd::launch_rk_stage_plan(P, F, u2, u, u, fixed, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0 * dt, s);
        }
        RBFFD_HIP(hipGetLastError());
    });
}

int rbffd_generate_operator(rbffd_context *c, const double *X, int64_t N, const double *Y, int64_t M, int32_t p,
                            int32_t n, int32_t polydeg, const rbffd_boundary *bnd, int32_t nops,
                            const int32_t *ops, const int32_t *opk, int32_t precision, int32_t index_base,
                            int64_t *colptr, int64_t *rowval, double *const *nzval, int64_t *ncols,
                            rbffd_stats *stats)
{
    return guarded([&] {
        check_ctx(c);
        if (!X || !ops || !colptr || !rowval || !nzval) throw Error{RBFFD_ERR_ARG, "null pointer"};
        check_ops(nops, ops, opk);
        if (N < 1) throw Error{RBFFD_ERR_ARG, "no X nodes"};
        const bool m3 = (Y == nullptr);
        const long Mr = m3 ? N : M;
        const long nnz = Mr * (long)n;
        if (nnz > INT_MAX || N > INT_MAX || Mr > INT_MAX)
            throw Error{RBFFD_ERR_ARG, "generate_operator: more than 2^31 - 1 entries or nodes"};
        hipStream_t s = c->stream;
        double *dX = (double *)ws(c, S_H_X, sizeof(double) * 2 * N);
        RBFFD_HIP(hipMemcpyAsync(dX, X, sizeof(double) * 2 * N, hipMemcpyHostToDevice, s));
        double *dY = nullptr;
        if (!m3) {
            if (Y == X) {
                dY = dX;
            } else {
                dY = (double *)ws(c, S_H_Y, sizeof(double) * 2 * Mr);
                RBFFD_HIP(hipMemcpyAsync(dY, Y, sizeof(double) * 2 * Mr, hipMemcpyHostToDevice, s));
            }
        }
        int *dcol = (int *)ws(c, S_H_COL, sizeof(int) * nnz);
        double *dv = (double *)ws(c, S_H_VALS, sizeof(double) * nnz * nops);
        double *dcv = (double *)ws(c, S_H_CVALS, sizeof(double) * nnz * nops);
        int64_t *dptr = (int64_t *)ws(c, S_H_PTR64, sizeof(int64_t) * (N + 1));
        int64_t *drow = (int64_t *)ws(c, S_H_ROW64, sizeof(int64_t) * nnz);
        double *vp[RBFFD_MAX_OPS], *cvp[RBFFD_MAX_OPS];
        for (int o = 0; o < nops; ++o) { vp[o] = dv + (size_t)o * nnz; cvp[o] = dcv + (size_t)o * nnz; }
        generate_impl(c, dX, N, dY, Mr, p, n, polydeg, bnd, nops, ops, opk, precision, dcol, vp, nullptr,
                      nullptr, stats, nullptr);
        int64_t used = 0;
        csc_impl(c, Mr, n, N, dcol, nops, vp, dptr, drow, cvp, index_base, &used);
        if (ncols) *ncols = used;
        RBFFD_HIP(hipMemcpyAsync(colptr, dptr, sizeof(int64_t) * (N + 1), hipMemcpyDeviceToHost, s));
        RBFFD_HIP(hipStreamSynchronize(s));
        d2h_staged(c, rowval, drow, sizeof(int64_t) * nnz);
        for (int o = 0; o < nops; ++o) d2h_staged(c, nzval[o], cvp[o], sizeof(double) * nnz);
    });
}

int rbffd_boundary_groups(const rbffd_boundary *bnd, int64_t N, int32_t *xgroup, int32_t *qgroup)
{
    return guarded([&] {
        if (!bnd || !xgroup || !qgroup || N < 1) throw Error{RBFFD_ERR_ARG, "boundary_groups: bad arguments"};
        std::vector<int> xg, qg;
        boundary_groups(bnd, N, xg, qg);
        std::copy(xg.begin(), xg.end(), xgroup);
        std::copy(qg.begin(), qg.end(), qgroup);
    });
}

int rbffd_generate_batch_device(rbffd_context *c, int32_t nsets, const rbffd_batch_set *sets, int32_t nops,
                                const int32_t *ops, const int32_t *opk, int32_t use_graph, int32_t *set_status,
                                rbffd_stats *stats)
{
    return guarded([&] { batch_impl(c, nsets, sets, nops, ops, opk, use_graph, set_status, stats); });
}

int rbffd_combine_device(rbffd_context *c, int64_t nnz, int32_t nops, const double *const *vals,
                         const double *coef, double *out, float *out32)
{
    return guarded([&] {
        check_ctx(c);
        if (nnz < 0 || nops < 1 || nops > rbffd::kMaxOps || !vals || !coef || (!out && !out32))
            throw Error{RBFFD_ERR_ARG, "combine: bad arguments"};
        rbffd::CombineTerms t{};
        t.nops = nops;
        for (int o = 0; o < nops; ++o) {
            if (!vals[o]) throw Error{RBFFD_ERR_ARG, "combine: null value array"};
            t.vals[o] = vals[o];
            t.coef[o] = coef[o];
        }
        rbffd::launch_combine(nnz, t, out, out32, c->stream);
        RBFFD_HIP(hipGetLastError());
    });
}

int rbffd_ssprk3_device(rbffd_context *c, int64_t N, const int64_t *rowptr, const int32_t *col, const void *vals,
                        int32_t vals_type, int32_t uniform_n, const int32_t *row_order, int32_t F, double *u,
                        double *work, const uint8_t *fixed, double dt, int64_t steps)
{
    return guarded([&] {
        check_ctx(c);
        if (N < 0 || !col || !vals || !u || !work || steps < 0) throw Error{RBFFD_ERR_ARG, "ssprk3: bad arguments"};
        if (uniform_n <= 0 && !rowptr) throw Error{RBFFD_ERR_ARG, "ssprk3: rowptr or uniform_n needed"};
        if (vals_type != RBFFD_F64 && vals_type != RBFFD_F32) throw Error{RBFFD_ERR_ARG, "ssprk3: value type"};
        if (F != 1 && F != 2 && F != 4 && F != 8) throw Error{RBFFD_ERR_UNSUPPORTED, "ssprk3: fields must be 1, 2, 4 or 8"};
        hipStream_t s = c->stream;
        const int f32 = vals_type == RBFFD_F32;
        double *u1 = work, *u2 = work + (size_t)N * F;
        for (int64_t k = 0; k < steps; ++k) {
            // Shu-Osher form of SSP-RK3 (Gottlieb, Shu and Tadmor 2001)
            const int *o = row_order;
            rbffd::launch_rk_stage(N, rowptr, col, vals, f32, F, uniform_n, o, u, u, u1, fixed, 0.0, 1.0, dt, s);
            rbffd::launch_rk_stage(N, rowptr, col, vals, f32, F, uniform_n, o, u1, u, u2, fixed, 0.75, 0.25,
                                   0.25 * dt, s);
            rbffd::launch_rk_stage(N, rowptr, col, vals, f32, F, uniform_n, o, u2, u, u, fixed, 1.0 / 3.0,
                                   2.0 / 3.0, 2.0 / 3.0 * dt, s);
        }
        RBFFD_HIP(hipGetLastError());
    });
}

int rbffd_dense_inverse_device(rbffd_context *c, int64_t n, const double *A, double *Ainv)
{
    return guarded([&] {
        check_ctx(c);
        if (n < 0 || (n > 0 && (!A || !Ainv))) throw Error{RBFFD_ERR_ARG, "inverse: bad arguments"};
        if (n == 0) return;
        if (n * n > (int64_t)INT_MAX) throw Error{RBFFD_ERR_ARG, "inverse: needs n^2 < 2^31"};
        double *Aw = (double *)ws(c, S_LSQ_DENSE, sizeof(double) * (size_t)n * n);
        int *info = (int *)ws(c, S_LSQ_INFO, sizeof(int) * 4);
        RBFFD_HIP(hipMemcpyAsync(Aw, A, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToDevice, c->stream));
        const int r = rbffd::dense_inverse(&c->solver, n, Aw, Ainv, info, c->stream);
        if (r > 0) throw Error{RBFFD_ERR_SINGULAR, "inverse: the matrix is singular"};
    });
}

int rbffd_ssprk3_closure_device(rbffd_context *c, int64_t N, const int64_t *rowptr, const int32_t *col,
                                const void *vals, int32_t vals_type, int32_t uniform_n, const int32_t *row_order,
                                int32_t F, double *u, double *work, const uint8_t *fixed, int32_t nclosures,
                                const rbffd_ghost_closure *closures, double dt, int64_t steps)
{
    return guarded([&] {
        check_ctx(c);
        if (N < 0 || !col || !vals || !u || !work || steps < 0) throw Error{RBFFD_ERR_ARG, "ssprk3: bad arguments"};
        if (uniform_n <= 0 && !rowptr) throw Error{RBFFD_ERR_ARG, "ssprk3: rowptr or uniform_n needed"};
        if (vals_type != RBFFD_F64 && vals_type != RBFFD_F32) throw Error{RBFFD_ERR_ARG, "ssprk3: value type"};
        if (F != 1 && F != 2 && F != 4 && F != 8) throw Error{RBFFD_ERR_UNSUPPORTED, "ssprk3: fields must be 1, 2, 4 or 8"};
        if (nclosures < 0 || (nclosures > 0 && !closures)) throw Error{RBFFD_ERR_ARG, "ssprk3: bad closures"};
        long maxb = 0;
        for (int m = 0; m < nclosures; ++m) {
            const rbffd_ghost_closure &g = closures[m];
            if (g.nbc < 1 || !g.rowptr || !g.col || !g.vals || !g.ginv || !g.ghost)
                throw Error{RBFFD_ERR_ARG, "ssprk3: closure " + std::to_string(m) + " is incomplete"};
            if (g.nbc > INT_MAX / 8) throw Error{RBFFD_ERR_ARG, "ssprk3: closure too large"};
            validate_csr(c, g.nbc, g.rowptr, g.col, N, "ssprk3 closure");
            validate_indices(c, g.ghost, g.nbc, N, "ssprk3 closure ghosts");
            maxb = std::max<long>(maxb, g.nbc);
        }
        hipStream_t s = c->stream;
        double *t = (double *)ws(c, S_CLOSURE_T, sizeof(double) * (size_t)std::max<long>(maxb, 1) * F);
        double *part = (double *)ws(c, S_CLOSURE_PART,
                                    sizeof(double) * (size_t)std::max<long>(rbffd::closure_chunks(maxb), 1) *
                                        std::max<long>(maxb, 1) * F);
        // ghost values of every marker, in marker order, from the current v
        auto close = [&](double *v) {
            for (int m = 0; m < nclosures; ++m) {
                const rbffd_ghost_closure &g = closures[m];
                rbffd::launch_spmv(g.nbc, g.rowptr, g.col, g.vals, 0, F, v, t, 1.0, 0.0, 0, nullptr, s);
                rbffd::launch_closure(g.nbc, F, g.ginv, t, part, g.ghost, v, s);
            }
        };
        const int f32 = vals_type == RBFFD_F32;
        double *u1 = work, *u2 = work + (size_t)N * F;
        const int *o = row_order;
        close(u);
        for (int64_t k = 0; k < steps; ++k) {
            rbffd::launch_rk_stage(N, rowptr, col, vals, f32, F, uniform_n, o, u, u, u1, fixed, 0.0, 1.0, dt, s);
            close(u1);
            rbffd::launch_rk_stage(N, rowptr, col, vals, f32, F, uniform_n, o, u1, u, u2, fixed, 0.75, 0.25,
                                   0.25 * dt, s);
            close(u2);
            rbffd::launch_rk_stage(N, rowptr, col, vals, f32, F, uniform_n, o, u2, u, u, fixed, 1.0 / 3.0,
                                   2.0 / 3.0, 2.0 / 3.0 * dt, s);
            close(u);
        }
        RBFFD_HIP(hipGetLastError());
    });
}

int rbffd_read_sum_device(rbffd_context *c, const double *x, int64_t n, double *out)
{
    return guarded([&] {
        check_ctx(c);
        if (!x || !out || n < 0) throw Error{RBFFD_ERR_ARG, "read_sum: bad arguments"};
        if (((uintptr_t)x & 15) != 0) throw Error{RBFFD_ERR_ARG, "read_sum: x must be 16-byte aligned"};
        int cus = 0;
        RBFFD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
        rbffd::launch_read_sum(x, n, out, cus, c->stream);
        RBFFD_HIP(hipGetLastError());
    });
}

int rbffd_locality_order_device(rbffd_context *c, const double *X, int64_t N, int32_t *order)
{
    return guarded([&] {
        check_ctx(c);
        if (!X || !order || N < 1) throw Error{RBFFD_ERR_ARG, "locality_order: bad arguments"};
        hipStream_t s = c->stream;
        const int nb = (int)std::min<long>(1024, (N + 255) / 256);
        double *part = (double *)ws(c, S_BBOX_PART, sizeof(double) * 5 * nb);
        double *box = (double *)ws(c, S_BBOX, sizeof(double) * 5);
        rbffd::launch_bbox(X, N, box, part, nb, s);
        auto *key = (uint32_t *)ws(c, S_MORTON_KEY, sizeof(uint32_t) * 2 * N);
        int *idx = (int *)ws(c, S_MORTON_IDX, sizeof(int) * N);
        const size_t tb = rbffd::morton_tmp_bytes(N);
        void *tmp = ws(c, S_SCAN_TMP2, tb);
        rbffd::launch_morton_order(X, N, box, key, key + N, idx, order, tmp, tb, s);
        RBFFD_HIP(hipGetLastError());
    });
}

}  // extern "C"
