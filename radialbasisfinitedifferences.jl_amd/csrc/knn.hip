// knn.hip -- K1: exact k-nearest-neighbour stencil selection on gfx950.
//
// Replaces the KDTree/knn calls of the reference (src/generate_operator.jl:43-47,
// src/hyperviscosity_operator.jl:54-58, src/calculateneighbors.jl:16-94).
// Points are bucketed into a uniform grid (counting sort by cell, cell-ordered
// copies of the coordinates so a cell scan is one contiguous read); one thread
// per query walks rings of cells outwards, skips cells farther than its current
// k-th distance and stops once the unvisited region is provably farther.
// The k best are kept sorted in registers.  Order is ascending squared
// distance dx*dx + dy*dy (never fused), ties by lower index, so the result
// does not depend on the order in which cells or points are visited.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <limits.h>

#include "device_common.hpp"
#include "internal.hpp"

namespace rbffd {

// ---------------------------------------------------------------- grid build
// Bounding box and the count of non-finite coordinates, 5 values per block.
__global__ void bbox_partial_kernel(const double *X, long N, double *partials)
{
    __shared__ double sm[5][256];
    double a = INFINITY, b = INFINITY, c = -INFINITY, d = -INFINITY, bad = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
        double x = X[2 * i], y = X[2 * i + 1];
        if (!__builtin_isfinite(x) || !__builtin_isfinite(y)) bad += 1.0;
        a = fmin(a, x); b = fmin(b, y); c = fmax(c, x); d = fmax(d, y);
    }
    sm[0][threadIdx.x] = a; sm[1][threadIdx.x] = b; sm[2][threadIdx.x] = c; sm[3][threadIdx.x] = d;
    sm[4][threadIdx.x] = bad;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sm[0][threadIdx.x] = fmin(sm[0][threadIdx.x], sm[0][threadIdx.x + o]);
            sm[1][threadIdx.x] = fmin(sm[1][threadIdx.x], sm[1][threadIdx.x + o]);
            sm[2][threadIdx.x] = fmax(sm[2][threadIdx.x], sm[2][threadIdx.x + o]);
            sm[3][threadIdx.x] = fmax(sm[3][threadIdx.x], sm[3][threadIdx.x + o]);
            sm[4][threadIdx.x] += sm[4][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) partials[blockIdx.x * 5 + threadIdx.x] = sm[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void bbox_final_kernel(const double *partials, int nblocks, double *box)
{
    __shared__ double sm[5][256];
    double a = INFINITY, b = INFINITY, c = -INFINITY, d = -INFINITY, bad = 0.0;
    for (int i = threadIdx.x; i < nblocks; i += blockDim.x) {
        a = fmin(a, partials[5 * i]); b = fmin(b, partials[5 * i + 1]);
        c = fmax(c, partials[5 * i + 2]); d = fmax(d, partials[5 * i + 3]);
        bad += partials[5 * i + 4];
    }
    sm[0][threadIdx.x] = a; sm[1][threadIdx.x] = b; sm[2][threadIdx.x] = c; sm[3][threadIdx.x] = d;
    sm[4][threadIdx.x] = bad;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sm[0][threadIdx.x] = fmin(sm[0][threadIdx.x], sm[0][threadIdx.x + o]);
            sm[1][threadIdx.x] = fmin(sm[1][threadIdx.x], sm[1][threadIdx.x + o]);
            sm[2][threadIdx.x] = fmax(sm[2][threadIdx.x], sm[2][threadIdx.x + o]);
            sm[3][threadIdx.x] = fmax(sm[3][threadIdx.x], sm[3][threadIdx.x + o]);
            sm[4][threadIdx.x] += sm[4][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) box[threadIdx.x] = sm[threadIdx.x][0];
}

void launch_bbox(const double *X, long N, double *box, double *partials, int nblocks, hipStream_t s)
{
    bbox_partial_kernel<<<nblocks, 256, 0, s>>>(X, N, partials);
    bbox_final_kernel<<<1, 256, 0, s>>>(partials, nblocks, box);
}

__device__ __forceinline__ int cell_coord(double v, double v0, double cs, int g)
{
    int c = (int)((v - v0) / cs);
    c = c < 0 ? 0 : c;
    return c >= g ? g - 1 : c;
}

// The count's atomic also hands each point its slot inside the cell, so the
// scatter needs neither a second atomic nor a zeroed fill array.
__global__ void cell_count_kernel(const double *X, long N, const GridParams *__restrict__ gp, int *cell_of,
                                  int *counts, int *slot)
{
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= N) return;
    const GridParams g = *gp;
    int cx = cell_coord(X[2 * i], g.x0, g.csx, g.gx);
    int cy = cell_coord(X[2 * i + 1], g.y0, g.csy, g.gy);
    int c = cy * g.gx + cx;
    cell_of[i] = c;
    slot[i] = atomicAdd(&counts[c], 1);
}

void launch_cell_count(const double *X, long N, const GridParams *gp, int *cell_of, int *counts, int *slot,
                       hipStream_t s)
{
    if (N <= 0) return;
    cell_count_kernel<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(X, N, gp, cell_of, counts, slot);
}

// Grid of about ppc points per cell over the bounding box, at most
// grid_max_cells(N) cells.  ppc = k/11, at least 2 (2.8 at k = 31): the
// selection kernel walks a row of cells as one range, so the cost of a cell
// is small and smaller cells leave fewer candidates outside the k-th distance
// (at k = 31 about 100 candidates a query where 5.2 points a cell gave 150).
// A box with non-finite entries (non-finite
// coordinates) gets a one-cell grid, which every later kernel handles safely,
// and sets bit 0 of *flag.
__global__ void grid_params_kernel(const double *box, long N, int k, GridParams *out, int *flag)
{
    const double x0 = box[0], y0 = box[1], w = box[2] - box[0], h = box[3] - box[1];
    GridParams g;
    if (box[4] != 0.0 || !__builtin_isfinite(w) || !__builtin_isfinite(h) || !__builtin_isfinite(x0) ||
        !__builtin_isfinite(y0)) {
        g.x0 = 0.0; g.y0 = 0.0; g.csx = 1.0; g.csy = 1.0; g.gx = 1; g.gy = 1; g.valid = 0;
        *out = g;
        if (flag) atomicOr(flag, 1);
        return;
    }
    const double ppc = fmax(2.0, k / 11.0);
    double cs;
    if (w > 0 && h > 0) cs = sqrt(w * h * ppc / (double)N);
    else cs = fmax(w, h) * ppc / (double)N;
    if (!(cs > 0)) cs = 1.0;
    const long cap = grid_max_cells(N);
    long gx, gy;
    for (;;) {
        gx = w > 0 ? min((long)(w / cs) + 1, 1L << 16) : 1;
        gy = h > 0 ? min((long)(h / cs) + 1, 1L << 16) : 1;
        if (gx * gy <= cap) break;
        cs *= 1.25;
    }
    g.x0 = x0;
    g.y0 = y0;
    g.gx = (int)gx;
    g.gy = (int)gy;
    g.csx = w > 0 ? w / (double)gx * (1.0 + 1e-12) : 1.0;
    g.csy = h > 0 ? h / (double)gy * (1.0 + 1e-12) : 1.0;
    g.valid = 1;
    *out = g;
}

void launch_grid_params(const double *box, long N, int k, GridParams *out, int *flag, hipStream_t s)
{
    grid_params_kernel<<<1, 1, 0, s>>>(box, N, k, out, flag);
}

__global__ void fill_i32_kernel(int *p, long n, int v)
{
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_fill_i32(int *p, long n, int v, hipStream_t s)
{
    if (n <= 0) return;
    fill_i32_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, n, v);
}

size_t scan_tmp_bytes(long n)
{
    size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const int *)nullptr, (int *)nullptr, (int)n);
    return bytes;
}

void launch_exclusive_scan(const int *in, int *out, long n, void *tmp, size_t tmp_bytes, hipStream_t s)
{
    RBFFD_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, in, out, (int)n, s));
}

__global__ void cell_scatter_kernel(const double *X, const int *xgroup, long N, const int *cell_of,
                                    const int *start, const int *slot, double *cell_xy, int *cell_idx,
                                    int *cell_grp)
{
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int pos = start[cell_of[i]] + slot[i];
    cell_xy[2 * (long)pos] = X[2 * i];
    cell_xy[2 * (long)pos + 1] = X[2 * i + 1];
    cell_idx[pos] = (int)i;
    if (cell_grp) cell_grp[pos] = xgroup[i];
}

void launch_cell_scatter(const double *X, const int *xgroup, long N, const int *cell_of, const int *start,
                         const int *slot, double *cell_xy, int *cell_idx, int *cell_grp, hipStream_t s)
{
    if (N <= 0) return;
    cell_scatter_kernel<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(X, xgroup, N, cell_of, start, slot,
                                                                     cell_xy, cell_idx, cell_grp);
}

// ---------------------------------------------------------------- query
__device__ __forceinline__ bool key_less(double d, int i, double d2, int i2)
{
    return d < d2 || (d == d2 && i < i2);
}

// Slots [0, KMAX-k) hold a sentinel smaller than any key, so the k real
// entries always end at slot KMAX-1 and every register index is static.
template <int KMAX>
__global__ __launch_bounds__(256) void knn_query_kernel(const GridParams *__restrict__ gp,
                                                        const int *__restrict__ cell_start,
                                                        const double *__restrict__ cell_xy,
                                                        const int *__restrict__ cell_idx,
                                                        const int *__restrict__ cell_grp,
                                                        const double *__restrict__ Q,
                                                        const int *__restrict__ qgroup,
                                                        const int *__restrict__ qorder, long M,
                                                        const int *__restrict__ mcount, int k,
                                                        int *__restrict__ idx_out, double *__restrict__ d2_out,
                                                        int *__restrict__ unfilled)
{
#pragma clang fp contract(off)
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (mcount ? (long)*mcount : M)) return;
    const GridParams g = *gp;
    const long q = qorder ? (long)qorder[t] : t;
    if (!g.valid) {  // non-finite points: no search, every slot -1
        for (int j = 0; j < k; ++j) {
            idx_out[q * k + j] = -1;
            if (d2_out) d2_out[q * k + j] = INFINITY;
        }
        if (unfilled) atomicAdd(unfilled, k);
        return;
    }
    const double qx = Q[2 * q], qy = Q[2 * q + 1];
    const int qg = (cell_grp && qgroup) ? qgroup[q] : -1;

    double kd[KMAX];
    int ki[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
        bool sentinel = j < KMAX - k;
        kd[j] = sentinel ? -1.0 : INFINITY;
        ki[j] = -1;
    }

    const double fx = (qx - g.x0) / g.csx, fy = (qy - g.y0) / g.csy;
    // clamp far-away queries into a range where int arithmetic is safe
    const int qcx = (int)fmin(fmax(floor(fx), -2.0), (double)g.gx + 1.0);
    const int qcy = (int)fmin(fmax(floor(fy), -2.0), (double)g.gy + 1.0);
    const int rlimit = g.gx + g.gy + 4;

    for (int r = 0; r <= rlimit; ++r) {
        const int ncell = r == 0 ? 1 : 8 * r;
        for (int tc = 0; tc < ncell; ++tc) {
            int dxc = 0, dyc = 0;
            if (r > 0) {
                int side = tc / (2 * r), off = tc - side * 2 * r;
                if (side == 0) { dxc = -r + off; dyc = -r; }
                else if (side == 1) { dxc = r; dyc = -r + off; }
                else if (side == 2) { dxc = r - off; dyc = r; }
                else { dxc = -r; dyc = r - off; }
            }
            const int cx = qcx + dxc, cy = qcy + dyc;
            if (cx < 0 || cx >= g.gx || cy < 0 || cy >= g.gy) continue;
            const double rx0 = g.x0 + cx * g.csx, rx1 = rx0 + g.csx;
            const double ry0 = g.y0 + cy * g.csy, ry1 = ry0 + g.csy;
            const double ddx = qx < rx0 ? rx0 - qx : (qx > rx1 ? qx - rx1 : 0.0);
            const double ddy = qy < ry0 ? ry0 - qy : (qy > ry1 ? qy - ry1 : 0.0);
            const double md = (ddx * ddx + ddy * ddy) * (1.0 - 1e-9);
            if (md > kd[KMAX - 1]) continue;
            const int c = cy * g.gx + cx;
            const int p0 = cell_start[c], p1 = cell_start[c + 1];
            for (int p = p0; p < p1; ++p) {
                const double px = cell_xy[2 * (long)p], py = cell_xy[2 * (long)p + 1];
                const int id = cell_idx[p];
                if (qg >= 0) {
                    const int h = cell_grp[p];
                    if (h != 0 && h != qg) continue;
                }
                const double dx = px - qx, dy = py - qy;
                const double dd = dx * dx;
                const double ee = dy * dy;
                const double d = dd + ee;
                if (key_less(d, id, kd[KMAX - 1], ki[KMAX - 1])) {
                    // sorted insertion; the last entry drops out
                    bool lt_next = true;
#pragma unroll
                    for (int j = KMAX - 1; j >= 1; --j) {
                        const bool lt = key_less(d, id, kd[j - 1], ki[j - 1]);
                        const double nd = lt ? kd[j - 1] : (lt_next ? d : kd[j]);
                        const int ni = lt ? ki[j - 1] : (lt_next ? id : ki[j]);
                        kd[j] = nd;
                        ki[j] = ni;
                        lt_next = lt;
                    }
                    if (lt_next) { kd[0] = d; ki[0] = id; }
                }
            }
        }
        // every unvisited cell lies outside the block of rings <= r
        const double bx0 = qx - (g.x0 + (double)(qcx - r) * g.csx);
        const double bx1 = (g.x0 + (double)(qcx + r + 1) * g.csx) - qx;
        const double by0 = qy - (g.y0 + (double)(qcy - r) * g.csy);
        const double by1 = (g.y0 + (double)(qcy + r + 1) * g.csy) - qy;
        const double b = fmin(fmin(bx0, bx1), fmin(by0, by1));
        if (b > 0.0 && b * b * (1.0 - 1e-9) > kd[KMAX - 1]) break;
        if (qcx - r <= 0 && qcy - r <= 0 && qcx + r >= g.gx - 1 && qcy + r >= g.gy - 1) break;
    }

    int missing = 0;
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
        if (j >= KMAX - k) {
            idx_out[q * k + (j - (KMAX - k))] = ki[j];
            if (d2_out) d2_out[q * k + (j - (KMAX - k))] = kd[j];
            missing += ki[j] < 0 ? 1 : 0;
        }
    }
    if (unfilled && missing) atomicAdd(unfilled, missing);
}

// Bits of a staged entry that hold the cell position when its rank (below
// KMAX) shares the entry: 26 up to KMAX = 32, 25 up to 64.
template <int KMAX>
constexpr int knn_pos_bits() { return KMAX <= 32 ? 26 : 25; }

// kv = min(kv, c), c = max(kv, c): two instructions.  __builtin_fmin/fmax
// first re-quiet both operands for the IEEE mode kernels run in, a third
// instruction a slot; no NaN reaches this chain (a NaN distance fails the
// d < kv test in front of it), so the plain instructions give the same values.
__device__ __forceinline__ void minmax_slot(double &kv, double &c)
{
    double hi;
    asm("v_max_f64 %0, %1, %2\n\tv_min_f64 %1, %1, %2" : "=&v"(hi), "+v"(kv) : "v"(c));
    c = hi;
}

// Two-phase exact selection for larger k.  Phase 1 keeps only the KMAX
// smallest squared distances, without their indices, in a min/max exchange
// chain (two instructions a slot where the insertion needs about nine), which
// yields tau, the k-th smallest.  Phase 2 walks the same cells again and
// writes every candidate with d <= tau at its rank among the kept values.
// That rank is exact when the kept values are distinct and exactly k
// candidates have d <= tau; any other query (an exact distance tie, fewer than
// k visible nodes, non-finite points) goes to `redo`, which the insertion
// kernel then recomputes with the (d^2, index) order.
//
// With STAGE, phase 2 places each candidate's cell-order position in an LDS
// row per thread (blockDim.x * k ints of dynamic shared memory) and each
// wavefront then writes its 64 rows out together, consecutive lanes on
// consecutive entries of a row, looking up the node index (and recomputing
// d^2 the same way when it is asked for): stores made straight from phase 2
// put 4 bytes into a different row per lane, spread over the whole cell walk,
// and the partly written lines cost about 2.7x the output size in HBM writes.
// Below 2^26 points (2^25 from KMAX = 48) the rows are staged in the order
// found and ranked after the walk, when every lane has its k candidates:
// during the walk about a third of the lanes accept a given candidate, and
// the rank costs the wavefront the same either way.  The rank is kept above
// the position bits of the staged entry, since storing each entry at its rank
// straight to HBM from the owning lane is as scattered as storing during the
// walk.  Rows sent to `redo` are not written here; the redo kernels write all
// k entries of those.
template <int KMAX, bool STAGE>
__global__ __launch_bounds__(256) void knn_select_kernel(const GridParams *__restrict__ gp,
                                                         const int *__restrict__ cell_start,
                                                         const double *__restrict__ cell_xy,
                                                         const int *__restrict__ cell_idx,
                                                         const int *__restrict__ cell_grp,
                                                         const double *__restrict__ Q,
                                                         const int *__restrict__ qgroup,
                                                         const int *__restrict__ qorder, long M, int k,
                                                         int *__restrict__ idx_out, double *__restrict__ d2_out,
                                                         int *__restrict__ redo, double *__restrict__ redo_tau,
                                                         int *__restrict__ redo_count, int pack)
{
#pragma clang fp contract(off)
    extern __shared__ int srow[];  // STAGE: [blockDim.x][k]
    __shared__ int srow_q[STAGE ? 256 : 1];
    const long t0 = blockIdx.x * (long)blockDim.x;
    const long t = t0 + threadIdx.x;
    int *const mine = STAGE ? srow + threadIdx.x * k : nullptr;
    // rank after the walk when a staged entry can hold a cell position (PB
    // bits: cell positions below 2^PB, which pack states) and its rank above it
    constexpr bool DENSE_OK = STAGE;
    constexpr int PB = knn_pos_bits<KMAX>();
    const bool dense = DENSE_OK && pack;
    // the query's row when every entry of it is placed, else -1
    const long qdone = [&]() -> long {
        if (t >= M) return -1;
        const GridParams g = *gp;
        const long q = qorder ? (long)qorder[t] : t;
        if (!g.valid) {
            const int slot = atomicAdd(redo_count, 1);
            redo[slot] = (int)q;
            redo_tau[slot] = INFINITY;
            return -1;
        }
        const double qx = Q[2 * q], qy = Q[2 * q + 1];
        const int qg = (cell_grp && qgroup) ? qgroup[q] : -1;

        // kept values ascending; slots [0, KMAX-k) hold a sentinel below any distance
        double kv[KMAX];
#pragma unroll
        for (int j = 0; j < KMAX; ++j) kv[j] = j < KMAX - k ? -1.0 : INFINITY;

        const double fx = (qx - g.x0) / g.csx, fy = (qy - g.y0) / g.csy;
        const int qcx = (int)fmin(fmax(floor(fx), -2.0), (double)g.gx + 1.0);
        const int qcy = (int)fmin(fmax(floor(fy), -2.0), (double)g.gy + 1.0);
        const int rlimit = g.gx + g.gy + 4;

        // The cells of ring r as runs along a grid row: the cells of one row
        // are contiguous in cell order, so a run is one range of points, and a
        // wavefront walks the 3x3 core as three runs of three cells where it
        // would otherwise set up nine cells, most holding fewer points than the
        // fullest cell of the same step on another lane.  Step 0 is rings 0
        // and 1 together; step r >= 2 is ring r: its top and bottom rows, then
        // the side cells.  f(first cell, one past the last cell, lower bound
        // on d^2 over the run).
        auto ring = [&](int r, auto &&f) {
            const int nrun = r == 0 ? 3 : 4 * r;
            for (int s = 0; s < nrun; ++s) {
                int a, b, dyc;
                if (r == 0) { a = -1; b = 1; dyc = s - 1; }
                else if (s < 2) { a = -r; b = r; dyc = s == 0 ? -r : r; }
                else { a = b = ((s - 2) & 1) ? r : -r; dyc = -r + 1 + ((s - 2) >> 1); }
                int cx0 = qcx + a, cx1 = qcx + b;
                const int cy = qcy + dyc;
                if (cx1 < 0 || cx0 >= g.gx || cy < 0 || cy >= g.gy) continue;
                cx0 = max(cx0, 0);
                cx1 = min(cx1, g.gx - 1);
                const double rx0 = g.x0 + cx0 * g.csx, rx1 = g.x0 + (cx1 + 1) * g.csx;
                const double ry0 = g.y0 + cy * g.csy, ry1 = ry0 + g.csy;
                // the gap to the run per axis (0 inside), without branches
                const double ddx = fmax(fmax(rx0 - qx, qx - rx1), 0.0);
                const double ddy = fmax(fmax(ry0 - qy, qy - ry1), 0.0);
                f(cy * g.gx + cx0, cy * g.gx + cx1 + 1, (ddx * ddx + ddy * ddy) * (1.0 - 1e-9));
            }
        };
        // squared distance to candidate p, or -1 when the query does not see it
        auto d2of = [&](int p) {
            const double dx = cell_xy[2 * (long)p] - qx, dy = cell_xy[2 * (long)p + 1] - qy;
            const double dd = dx * dx;
            const double ee = dy * dy;
            return dd + ee;
        };
        auto dist = [&](int p) {
            if (qg >= 0) {
                const int h = cell_grp[p];
                if (h != 0 && h != qg) return -1.0;
            }
            return d2of(p);
        };

        // ---- phase 1: the KMAX smallest values
        int rlast = 0;
        for (int r = 0; r <= rlimit; r = r ? r + 1 : 2) {
            rlast = r;
            ring(r, [&](int c0, int c1, double md) {
                if (md > kv[KMAX - 1]) return;
                const int p1 = cell_start[c1];
                for (int p = cell_start[c0]; p < p1; ++p) {
                    const double d = dist(p);
                    if (d >= 0.0 && d < kv[KMAX - 1]) {
                        double cv = d;
#pragma unroll
                        for (int j = 0; j < KMAX; ++j) minmax_slot(kv[j], cv);
                    }
                }
            });
            // every ring up to rb is walked
            const int rb = r ? r : 1;
            const double bx0 = qx - (g.x0 + (double)(qcx - rb) * g.csx);
            const double bx1 = (g.x0 + (double)(qcx + rb + 1) * g.csx) - qx;
            const double by0 = qy - (g.y0 + (double)(qcy - rb) * g.csy);
            const double by1 = (g.y0 + (double)(qcy + rb + 1) * g.csy) - qy;
            const double b = fmin(fmin(bx0, bx1), fmin(by0, by1));
            if (b > 0.0 && b * b * (1.0 - 1e-9) > kv[KMAX - 1]) break;
            if (qcx - rb <= 0 && qcy - rb <= 0 && qcx + rb >= g.gx - 1 && qcy + rb >= g.gy - 1) break;
        }

        // ---- phase 2: every candidate with d <= tau, at its rank
        const double tau = kv[KMAX - 1];
        bool tie = !(tau < INFINITY);
#pragma unroll
        for (int j = 1; j < KMAX; ++j)
            if (j > KMAX - k) tie |= kv[j] == kv[j - 1];
        // rank of d among the kept values: a compare and an add with carry a
        // slot, kept apart from the address arithmetic, which the compiler
        // would otherwise fold into every term
        auto rank = [&](double d) {
            int pos = 0;
#pragma unroll
            for (int j = 0; j < KMAX; ++j) pos += kv[j] < d ? 1 : 0;
            asm volatile("" : "+v"(pos));
            return pos - (KMAX - k);
        };
        int nsel = 0;
        if (!tie) {
            for (int r = 0; r <= rlast; r = r ? r + 1 : 2) {
                ring(r, [&](int c0, int c1, double md) {
                    if (md > tau) return;
                    const int p1 = cell_start[c1];
                    for (int p = cell_start[c0]; p < p1; ++p) {
                        const double d = dist(p);
                        if (d >= 0.0 && d <= tau) {
                            if (dense) {
                                if (nsel < k) mine[nsel] = p;
                            } else {
                                const int pos = rank(d);
                                if (STAGE) {
                                    mine[pos] = p;
                                } else {
                                    idx_out[q * k + pos] = cell_idx[p];
                                    if (d2_out) d2_out[q * k + pos] = d;
                                }
                            }
                            ++nsel;
                        }
                    }
                });
            }
        }
        if (tie || nsel != k) {
            const int slot = atomicAdd(redo_count, 1);
            redo[slot] = (int)q;
            redo_tau[slot] = tie ? tau : (nsel > k ? tau : INFINITY);
            return -1;
        }
        if constexpr (DENSE_OK) {
            if (dense) {
                // the k candidates, staged in the order found, ranked with
                // every lane of the wavefront at work (four at a time, so their
                // coordinate loads overlap); the rank goes above bit PB
                int j = 0;
                for (; j + 4 <= k; j += 4) {
                    int p[4];
                    double d[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) p[u] = mine[j + u];
#pragma unroll
                    for (int u = 0; u < 4; ++u) d[u] = d2of(p[u]);
#pragma unroll
                    for (int u = 0; u < 4; ++u) mine[j + u] = p[u] | rank(d[u]) << PB;
                }
                for (; j < k; ++j) mine[j] = mine[j] | rank(d2of(mine[j])) << PB;
            }
        }
        return q;
    }();
    if (!STAGE) return;

    // ---- the wavefront's rows, written out entry by entry (a wavefront reads
    // only the rows its own lanes staged, so no other wavefront is waited for)
    const int lane = (int)threadIdx.x & 63, wbase = (int)threadIdx.x & ~63;
    srow_q[threadIdx.x] = (int)qdone;
    __builtin_amdgcn_wave_barrier();
    const long w0 = t0 + wbase;
    const int nw = (int)(M - w0 < 64 ? M - w0 : 64);
    const int *const wrow = srow + wbase * k;
    const int step_r = 64 / k, step_p = 64 - step_r * k;
    int r = lane / k, pos = lane - r * k;
    for (int e = lane; e < nw * k; e += 64) {
        const int qq = srow_q[wbase + r];
        if (qq >= 0) {
            int p = wrow[e], at = pos;
            if (dense) {
                at = p >> PB;
                p &= (1 << PB) - 1;
            }
            idx_out[(long)qq * k + at] = cell_idx[p];
            if (d2_out) {
                const double dx = cell_xy[2 * (long)p] - Q[2 * (long)qq];
                const double dy = cell_xy[2 * (long)p + 1] - Q[2 * (long)qq + 1];
                const double dd = dx * dx;
                const double ee = dy * dy;
                d2_out[(long)qq * k + at] = dd + ee;
            }
        }
        r += step_r;
        pos += step_p;
        if (pos >= k) {
            pos -= k;
            ++r;
        }
    }
}

// Queries the two-phase selection could not place (an exact distance tie, or
// more than k candidates at d <= tau): one wavefront per query.  tau, the
// exact k-th smallest distance from phase 1, bounds the cell walk as it did
// there; the lanes split each ring's cells and append every visible candidate
// with d <= tau to an LDS list, and each candidate's rank in the (d^2, index)
// order is counted against the list.  The first k ranks are the answer.
// Queries without a finite tau (fewer than k visible nodes, non-finite
// points) or with more than kTieCap candidates go on to the insertion kernel.
constexpr int kTieCap = 256;
constexpr int kTieWaves = 4;

__global__ __launch_bounds__(64 * kTieWaves) void knn_tie_kernel(
    const GridParams *__restrict__ gp, const int *__restrict__ cell_start, const double *__restrict__ cell_xy,
    const int *__restrict__ cell_idx, const int *__restrict__ cell_grp, const double *__restrict__ Q,
    const int *__restrict__ qgroup, const int *__restrict__ redo, const double *__restrict__ redo_tau,
    const int *__restrict__ redo_count, int k, int *__restrict__ idx_out, double *__restrict__ d2_out,
    int *__restrict__ redo2, int *__restrict__ redo2_count)
{
#pragma clang fp contract(off)
    __shared__ double cd[kTieWaves][kTieCap];
    __shared__ int ci[kTieWaves][kTieCap];
    __shared__ int cnt[kTieWaves];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const GridParams g = *gp;
    const long nred = *redo_count;
    const long stride = (long)gridDim.x * kTieWaves;
    for (long e = (long)blockIdx.x * kTieWaves + w; e < nred; e += stride) {  // uniform per wavefront
        const long q = redo[e];
        const double tau = redo_tau[e];
        if (!g.valid || !(tau < INFINITY)) {
            if (lane == 0) redo2[atomicAdd(redo2_count, 1)] = (int)q;
            continue;
        }
        const double qx = Q[2 * q], qy = Q[2 * q + 1];
        const int qg = (cell_grp && qgroup) ? qgroup[q] : -1;
        const double fx = (qx - g.x0) / g.csx, fy = (qy - g.y0) / g.csy;
        const int qcx = (int)fmin(fmax(floor(fx), -2.0), (double)g.gx + 1.0);
        const int qcy = (int)fmin(fmax(floor(fy), -2.0), (double)g.gy + 1.0);
        const int rlimit = g.gx + g.gy + 4;
        if (lane == 0) cnt[w] = 0;
        __builtin_amdgcn_wave_barrier();
        for (int r = 0; r <= rlimit; ++r) {
            const int ncell = r == 0 ? 1 : 8 * r;
            for (int tc = lane; tc < ncell; tc += 64) {
                int dxc = 0, dyc = 0;
                if (r > 0) {
                    const int side = tc / (2 * r), off = tc - side * 2 * r;
                    if (side == 0) { dxc = -r + off; dyc = -r; }
                    else if (side == 1) { dxc = r; dyc = -r + off; }
                    else if (side == 2) { dxc = r - off; dyc = r; }
                    else { dxc = -r; dyc = r - off; }
                }
                const int cx = qcx + dxc, cy = qcy + dyc;
                if (cx < 0 || cx >= g.gx || cy < 0 || cy >= g.gy) continue;
                const double rx0 = g.x0 + cx * g.csx, rx1 = rx0 + g.csx;
                const double ry0 = g.y0 + cy * g.csy, ry1 = ry0 + g.csy;
                const double ddx = qx < rx0 ? rx0 - qx : (qx > rx1 ? qx - rx1 : 0.0);
                const double ddy = qy < ry0 ? ry0 - qy : (qy > ry1 ? qy - ry1 : 0.0);
                if ((ddx * ddx + ddy * ddy) * (1.0 - 1e-9) > tau) continue;
                const int c = cy * g.gx + cx;
                const int p1 = cell_start[c + 1];
                for (int p = cell_start[c]; p < p1; ++p) {
                    if (qg >= 0) {
                        const int h = cell_grp[p];
                        if (h != 0 && h != qg) continue;
                    }
                    const double dx = cell_xy[2 * (long)p] - qx, dy = cell_xy[2 * (long)p + 1] - qy;
                    const double dd = dx * dx;
                    const double ee = dy * dy;
                    const double d = dd + ee;
                    if (d <= tau) {
                        const int slot = atomicAdd(&cnt[w], 1);
                        if (slot < kTieCap) {
                            cd[w][slot] = d;
                            ci[w][slot] = cell_idx[p];
                        }
                    }
                }
            }
            const double bx0 = qx - (g.x0 + (double)(qcx - r) * g.csx);
            const double bx1 = (g.x0 + (double)(qcx + r + 1) * g.csx) - qx;
            const double by0 = qy - (g.y0 + (double)(qcy - r) * g.csy);
            const double by1 = (g.y0 + (double)(qcy + r + 1) * g.csy) - qy;
            const double b = fmin(fmin(bx0, bx1), fmin(by0, by1));
            if (b > 0.0 && b * b * (1.0 - 1e-9) > tau) break;
            if (qcx - r <= 0 && qcy - r <= 0 && qcx + r >= g.gx - 1 && qcy + r >= g.gy - 1) break;
        }
        __builtin_amdgcn_wave_barrier();
        const int C = cnt[w];
        if (C > kTieCap || C < k) {
            if (lane == 0) redo2[atomicAdd(redo2_count, 1)] = (int)q;
            continue;
        }
        for (int i = lane; i < C; i += 64) {
            const double di = cd[w][i];
            const int ii = ci[w][i];
            int rank = 0;
            for (int j = 0; j < C; ++j) rank += key_less(cd[w][j], ci[w][j], di, ii) ? 1 : 0;
            if (rank < k) {
                idx_out[q * k + rank] = ii;
                if (d2_out) d2_out[q * k + rank] = di;
            }
        }
        // the next query's count starts at 0 only after every lane has read this one
        __builtin_amdgcn_wave_barrier();
    }
}

template <int KMAX>
static void knn_launch_t(const GridParams *g, const int *cs, const double *cxy, const int *cidx, const int *cgrp,
                         long N, const double *Q, const int *qgroup, const int *qorder, long M, int k, int *idx,
                         double *d2, int *unfilled, int *redo, int *redo_count, const KnnRedo &rw, hipStream_t s)
{
    const unsigned blocks = (unsigned)((M + 255) / 256);
    if (KMAX >= 16 && redo) {
        // rows staged in LDS up to 64 KiB a block (k <= 64)
        constexpr bool stage = KMAX <= 64;
        const size_t lds = stage ? sizeof(int) * 256 * (size_t)k : 0;
        knn_select_kernel<KMAX, stage><<<blocks, 256, lds, s>>>(g, cs, cxy, cidx, cgrp, Q, qgroup, qorder, M, k,
                                                                 idx, d2, redo, rw.tau, redo_count,
                                                                 N <= (1L << knn_pos_bits<KMAX>()) ? 1 : 0);
        // at most 16384 wavefronts walk the tied queries in turn
        const unsigned tblocks = (unsigned)std::max(1L, std::min<long>((M + kTieWaves - 1) / kTieWaves, 4096));
        knn_tie_kernel<<<tblocks, 64 * kTieWaves, 0, s>>>(g, cs, cxy, cidx, cgrp, Q, qgroup, redo, rw.tau,
                                                           redo_count, k, idx, d2, rw.list2, rw.count2);
        knn_query_kernel<KMAX><<<blocks, 256, 0, s>>>(g, cs, cxy, cidx, cgrp, Q, qgroup, rw.list2, M, rw.count2, k,
                                                       idx, d2, unfilled);
    } else {
        knn_query_kernel<KMAX><<<blocks, 256, 0, s>>>(g, cs, cxy, cidx, cgrp, Q, qgroup, qorder, M, nullptr, k,
                                                       idx, d2, unfilled);
    }
}

void launch_knn_query(const GridParams *g, const int *cs, const double *cxy, const int *cidx, const int *cgrp,
                      long N, const double *Q, const int *qgroup, const int *qorder, long M, int k, int *idx,
                      double *d2, int *unfilled, int *redo, int *redo_count, const KnnRedo &rw, hipStream_t s)
{
    if (M <= 0) return;
    if (redo) {
        launch_fill_i32(redo_count, 1, 0, s);
        launch_fill_i32(rw.count2, 1, 0, s);
    }
#define RBFFD_KNN(KM) \
    knn_launch_t<KM>(g, cs, cxy, cidx, cgrp, N, Q, qgroup, qorder, M, k, idx, d2, unfilled, redo, redo_count, rw, s)
    if (k <= 1) RBFFD_KNN(1);
    else if (k <= 8) RBFFD_KNN(8);
    else if (k <= 16) RBFFD_KNN(16);
    else if (k <= 24) RBFFD_KNN(24);
    else if (k <= 32) RBFFD_KNN(32);
    else if (k <= 48) RBFFD_KNN(48);
    else if (k <= 64) RBFFD_KNN(64);
    else if (k <= 96) RBFFD_KNN(96);
    else if (k <= 128) RBFFD_KNN(128);
    else throw Error{2, "knn: k > 128 is not supported"};
#undef RBFFD_KNN
}

}  // namespace rbffd
