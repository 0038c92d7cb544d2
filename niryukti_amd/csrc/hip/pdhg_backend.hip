// MI355X (gfx950) PDHG iteration backend for one LP.
//
// One PDHG attempt is two kernel launches:
//
//   primal_kernel : g = c + A^T y (CSR SpMV over the explicit transpose, one
//                   lane group per row), fused with the projected primal step
//                   x~ = clip(x - tau g, lb, ub), dx = x~ - x, the speculative
//                   running average and the per-block partial of |dx|^2.
//   dual_kernel   : delta = A dx (CSR SpMV), fused with the extrapolated dual
//                   prox step, the speculative activity/average updates and the
//                   partials of |dy|^2 and dy.delta; the last block to finish
//                   reduces all partials in a fixed order and takes the
//                   accept/reject and step-factor decision on the device.
//
// Iterates are double-buffered: kernels read buffer `cur` and write `cur ^ 1`;
// an accepted step flips `cur` in the device state, a rejected one leaves it,
// so no commit kernel and no host round trip is needed per attempt.  A chunk of
// attempts is launched back to back (or replayed as one hipGraph); every
// kernel first reads the device state and returns at once when the chunk has
// its accepted steps, so a chunk performs exactly `count` accepted steps as the
// reference CPU backend does (src/cpu.cpp:103-220) while the host synchronises
// once per chunk.  All reductions use fixed trees, so results are
// deterministic and independent of other work on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "niryukti_pdhg.h"

namespace {

thread_local std::string g_error;

#define HIP_OK(expr)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 2048;
constexpr int kSpareAttempts = 4;  // early-out attempts appended to each chunk for rejections
constexpr int kMaxRejects = 40;    // consecutive rejections before giving up (src/cpu.cpp:119)
constexpr int kInlineDecideBlocks = 256;  // larger dual grids decide in decide_kernel

enum Stop : int { kRun = 0, kDone = 1, kRestart = 2, kExhausted = 3, kNonfinite = 4 };

struct State {
    double tau0, sigma0;  // base steps of the current chunk
    double factor;        // adaptive step factor
    double average_mass;  // sum of accepted factors since the last restart
    double initial_residual, previous_residual;
    long long accepted, rejected, epoch_steps, chunk_accepted, chunk_target;
    int cur;  // which half of the double-buffered iterates is current
    int consecutive_rejects;
    int stop;
    int restart_requested;
    unsigned ticket;  // last-block-done counter of dual_kernel
    unsigned pad;
};

struct Flags {
    int adaptive, halpern, residual_restarts, pad;
    double reflection;
};

template <class O, class V>
struct DevCsr {
    const O *ptr;
    const int *idx;
    const V *val;
    long long rows;
};

// Rows longer than `threshold` nonzeros are not walked by a lane group: they
// are cut into segments of at most kSegment nonzeros, one workgroup each,
// placed at the start of the grid.  Every segment block stores its partial sum;
// the block that completes a row (the last to take the row's ticket) adds the
// row's partials in segment order and runs the row's epilogue, so the result
// does not depend on scheduling.  Power-law rows (tens of thousands of
// nonzeros) then spread over many CUs instead of serialising one lane group.
struct Segments {
    const long long *begin, *end;    // nonzero range of segment s
    const int *owner;                // long-row slot of segment s
    const int *row, *first, *count;  // per long-row slot: row, first segment, segments
    double *partial;                 // per segment
    unsigned *ticket;                // per long-row slot; zero between launches
    int nseg;
    long long threshold;
};

struct PrimalArgs {
    const double *c, *lb, *ub;
    double *x[2], *xa[2];
    double *dx;
    float *dx32;           // FP32 arithmetic phase: dx as gathered by A dx
    const double *y[2];
    const float *y32[2];   // FP32 arithmetic phase: the FP32 copy of y gathered by A^T y
    const double *anchor_x;
    double *partial;
    const double *part;  // chunk partials of A^T y, or null
    int parts;
    State *st;
    Flags f;
};

struct DualArgs {
    const double *rl, *ru;
    double *y[2], *ya[2], *ax[2];
    float *y32[2];          // FP32 arithmetic phase: FP32 copy of y, written beside y
    const double *dx;
    const float *dx32;
    const double *anchor_y, *anchor_ax;
    double *partial;  // 2 per block: |dy|^2, dy.delta
    const double *primal_partial;
    int primal_blocks;
    const double *part;  // chunk partials of A dx, or null
    int parts;
    int inline_decide;   // 1: last block decides; 0: decide_kernel runs next
    State *st;
    Flags f;
};

// ---------------------------------------------------------------- device --

template <int G, class T = double>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over a 256-thread block; the result is valid in thread 0.  Fixed order.
__device__ __forceinline__ double block_sum(double v, double *lds) {
    v = group_sum<64>(v);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0) r = (lds[0] + lds[1]) + (lds[2] + lds[3]);
    __syncthreads();
    return r;
}

__device__ __forceinline__ double clampd(double v, double lo, double hi) {  // std::clamp semantics
    return v < lo ? lo : (hi < v ? hi : v);
}
__device__ __forceinline__ double maxd(double a, double b) { return a < b ? b : a; }  // std::max
__device__ __forceinline__ double mind(double a, double b) { return b < a ? b : a; }  // std::min

__device__ __forceinline__ void store_shared(double *p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_shared(const double *p) {
    return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// This thread's fixed-order sum of values[stride * b + offset] over
// b = threadIdx.x, threadIdx.x + kBlock, ... < count, eight loads in flight.
__device__ __forceinline__ double gather_sum(const double *values, int stride, int offset, int count) {
    constexpr int U = 8;
    double sum = 0;
    for (int b0 = threadIdx.x; b0 < count; b0 += U * kBlock) {
        double t[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int b = b0 + u * kBlock;
            t[u] = b < count ? load_shared(&values[(long long)stride * b + offset]) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) sum += t[u];
    }
    return sum;
}

// Segment `blockIdx.x` of a long row: partial sum (and, with ABS, the sum of
// absolute products), ticket, and in thread 0 of the block that completes the
// row, the row index and its totals.
// T is the type of the gathered vector and of each thread's running sum:
// double, or float in the FP32 arithmetic phase (the partials across the
// threads of a block, and across segments, are added in double).
template <bool ABS, class O, class V, class T = double>
__device__ bool segment_row(const DevCsr<O, V> &A, const Segments &sg, const T *v, double *lds, long long &row,
                            double &total, double &total_abs, int s, int *completed = nullptr) {
    const long long end = sg.end[s];
    T acc = 0, acca = 0;
    for (long long k = sg.begin[s] + threadIdx.x; k < end; k += kBlock) {
        const T p = T(A.val[k]) * v[A.idx[k]];
        acc += p;
        if (ABS) acca += fabs(p);
    }
    const double part = block_sum(double(acc), lds);
    const double parta = ABS ? block_sum(double(acca), lds) : 0.0;
    __shared__ int completes;  // long-row slot this block completes, else -1
    if (threadIdx.x == 0) {
        store_shared(&sg.partial[2 * s], part);
        store_shared(&sg.partial[2 * s + 1], parta);
        const int L = sg.owner[s];
        const unsigned t = __hip_atomic_fetch_add(&sg.ticket[L], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        completes = t + 1 == unsigned(sg.count[L]) ? L : -1;
    }
    __syncthreads();
    const int L = completes;
    if (completed) *completed = L;
    if (L < 0) return false;
    // The whole block reads the row's segment partials at once (fixed order).
    const double *first = sg.partial + 2 * (long long)sg.first[L];
    const double sum = block_sum(gather_sum(first, 2, 0, sg.count[L]), lds);
    const double suma = ABS ? block_sum(gather_sum(first, 2, 1, sg.count[L]), lds) : 0.0;
    if (threadIdx.x == 0) {
        __hip_atomic_store(&sg.ticket[L], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        row = sg.row[L];
        total = sum;
        total_abs = suma;
    }
    return threadIdx.x == 0;
}

// Calls row(i, sum_k A[i,k] v[k], sum_k |A[i,k] v[k]|) for every row (the
// second sum only with ABS, else 0): long rows through their segments, short
// rows by lane groups of G over the remaining blocks.
//
// *slot (when given) receives where this block's share of any sum formed in
// row() belongs: blockIdx.x for row-group blocks; for segment blocks, the
// block of the row's first segment when this block completed a long row, and
// -1 otherwise.  Which segment block completes a row depends on timing, so
// writing the row's share under the completing block's own index would move
// it between partial slots from run to run and change the rounding of the
// fixed-order sum over slots.
template <int G, bool ABS = false, class O, class V, class T, class Row>
__device__ __forceinline__ void for_rows(const DevCsr<O, V> &A, const Segments &sg, const T *v, double *lds,
                                         Row &&row, long long *slot = nullptr) {
    if ((int)blockIdx.x < sg.nseg) {
        long long i = 0;
        double total = 0, total_abs = 0;
        int completed = -1;
        if (segment_row<ABS>(A, sg, v, lds, i, total, total_abs, blockIdx.x, &completed)) row(i, total, total_abs);
        if (slot) *slot = completed >= 0 ? (long long)sg.first[completed] : -1;
        return;
    }
    if (slot) *slot = blockIdx.x;
    constexpr int rows_per_block = kBlock / G;
    const int lane = threadIdx.x % G;
    const long long nb = (long long)gridDim.x - sg.nseg;
    for (long long base = ((long long)blockIdx.x - sg.nseg) * rows_per_block; base < A.rows;
         base += nb * rows_per_block) {
        const long long i = base + threadIdx.x / G;
        bool valid = i < A.rows;
        T acc = 0, acca = 0;
        if (valid) {
            const O begin = A.ptr[i], end = A.ptr[i + 1];
            valid = (long long)(end - begin) <= sg.threshold;
            if (valid)
                for (O k = begin + lane; k < end; k += G) {
                    const T p = T(A.val[k]) * v[A.idx[k]];
                    acc += p;
                    if (ABS) acca += fabs(p);
                }
        }
        acc = group_sum<G, T>(acc);
        if (ABS) acca = group_sum<G, T>(acca);
        if (valid && lane == 0) row(i, double(acc), double(acca));
    }
}

// XCD-aware split of a product y = M v whose gathered vector v is larger than
// one XCD's L2: the columns of M are cut into P contiguous chunks (P divides
// the 8 XCDs), chunk k stored as its own CSR over all rows.  Workgroups are
// dispatched to XCDs round-robin, so workgroup b runs on XCD b % 8 and works
// on chunk (b % 8) % P only: every XCD gathers from 1/P of v, which stays in
// its L2, instead of missing to the memory side for almost every gather.
// Partial row sums land in part[k * rows + i]; the consuming kernel adds the
// P partials in chunk order (a fixed order).  Rows handled by long-row
// segments are left out of the chunks.
struct ChunkedCsr {
    const int *ptr;     // P blocks of rows + 1 offsets, each relative to its chunk base
    const int *idx;     // global column indices, chunk after chunk
    long long base[8];  // start of chunk k in idx / val
    long long rows;
    int P;
};

template <class T>
struct VecPair {
    const T *p[2];  // the gathered vector by iterate parity (both the same when not double-buffered)
};

constexpr int kChunkRows = 4;  // rows per thread per pass of partial_kernel

// Partial sums of one column chunk, several rows per thread with the offsets,
// the first two entries of every row and their gathers issued together (rows
// average 1-2 nonzeros per chunk, so this is a latency chain, not a stream).
// Blocks [0, grid_chunks) work on the chunks; the blocks after them run the
// long-row segments of the whole matrix and write each long row's total into
// plane P, which is zero for every other row.
template <class O, class V, class T>
__global__ __launch_bounds__(kBlock) void partial_kernel(ChunkedCsr C, const V *val, DevCsr<O, V> A, Segments sg,
                                                         int grid_chunks, const State *st, VecPair<T> vec,
                                                         double *part) {
    if (st->stop != kRun) return;
    const T *v = vec.p[st->cur];
    if ((int)blockIdx.x >= grid_chunks) {
        __shared__ double lds[4];
        long long i = 0;
        double total = 0, total_abs = 0;
        if (segment_row<false>(A, sg, v, lds, i, total, total_abs, blockIdx.x - grid_chunks))
            part[(long long)C.P * C.rows + i] = total;
        return;
    }
    const int xcd = blockIdx.x & 7;
    const int k = xcd % C.P;
    const long long q = (long long)(blockIdx.x >> 3) * (8 / C.P) + xcd / C.P;
    const long long nq = grid_chunks / C.P;
    const int *ptr = C.ptr + (long long)k * (C.rows + 1);
    const int *idx = C.idx + C.base[k];
    const V *vals = val + C.base[k];
    double *out = part + (long long)k * C.rows;
    constexpr long long span = (long long)kBlock * kChunkRows;
    for (long long b = q * span; b < C.rows; b += nq * span) {
        int beg[kChunkRows], end[kChunkRows];
        double acc[kChunkRows];
#pragma unroll
        for (int r = 0; r < kChunkRows; ++r) {
            const long long i = b + threadIdx.x + (long long)r * kBlock;
            beg[r] = i < C.rows ? ptr[i] : 0;
            end[r] = i < C.rows ? ptr[i + 1] : 0;
        }
        int c0[kChunkRows], c1[kChunkRows];
        T v0[kChunkRows], v1[kChunkRows];
#pragma unroll
        for (int r = 0; r < kChunkRows; ++r) {
            const bool h0 = beg[r] < end[r], h1 = beg[r] + 1 < end[r];
            c0[r] = h0 ? idx[beg[r]] : 0;
            v0[r] = h0 ? T(vals[beg[r]]) : T(0);
            c1[r] = h1 ? idx[beg[r] + 1] : 0;
            v1[r] = h1 ? T(vals[beg[r] + 1]) : T(0);
        }
        T sum[kChunkRows];
#pragma unroll
        for (int r = 0; r < kChunkRows; ++r) {
            const T g0 = beg[r] < end[r] ? v[c0[r]] : T(0);
            const T g1 = beg[r] + 1 < end[r] ? v[c1[r]] : T(0);
            sum[r] = 0;
            if (beg[r] < end[r]) sum[r] += v0[r] * g0;
            if (beg[r] + 1 < end[r]) sum[r] += v1[r] * g1;
            for (int e = beg[r] + 2; e < end[r]; ++e) sum[r] += T(vals[e]) * v[idx[e]];
            acc[r] = double(sum[r]);
        }
#pragma unroll
        for (int r = 0; r < kChunkRows; ++r) {
            const long long i = b + threadIdx.x + (long long)r * kBlock;
            if (i < C.rows) out[i] = acc[r];
        }
    }
}

// for_rows with every row summed from the P + 1 planes of partial_kernel
// (fixed order), one thread per row: a pure stream.
template <class Row>
__device__ __forceinline__ void for_rows_partials(long long rows, const double *part, int P, Row &&row) {
    const long long stride = (long long)gridDim.x * kBlock;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < rows; i += stride) {
        double acc = 0;
        for (int k = 0; k <= P; ++k) acc += part[(long long)k * rows + i];
        row(i, acc, 0.0);
    }
}

// The vector a product gathers: the FP64 iterate, or in the FP32 arithmetic
// phase (T = float) its FP32 copy.
template <class T>
__device__ __forceinline__ const T *gathered(const double *d, const float *f);
template <>
__device__ __forceinline__ const double *gathered<double>(const double *d, const float *) {
    return d;
}
template <>
__device__ __forceinline__ const float *gathered<float>(const double *, const float *f) {
    return f;
}

// T: arithmetic of A^T y (double, or float in the FP32 arithmetic phase,
// where A^T y gathers an FP32 copy of y, sums in FP32 per lane and the
// projected step itself stays FP64).
template <int G, class O, class V, class T = double>
__global__ __launch_bounds__(kBlock) void primal_kernel(DevCsr<O, V> at, Segments sg, PrimalArgs a) {
    constexpr bool kSingle = sizeof(T) == 4;
    const State *st = a.st;
    if (st->stop != kRun) return;
    const int cur = st->cur;
    const T *y = gathered<T>(a.y[cur], a.y32[cur]);
    const double *x = a.x[cur];
    double *xn = a.x[cur ^ 1];
    const double *xa = a.xa[cur];
    double *xan = a.xa[cur ^ 1];
    const double f = st->factor;
    const double tau = st->tau0 * f;
    const double w = f / (st->average_mass + f);
    const double lambda = 1 / (double(st->epoch_steps) + 2);
    const double rho = a.f.reflection;
    double local = 0;
    __shared__ double lds[4];
    auto update = [&](long long j, double acc, double) {
        const double g = a.c[j] + acc;
        const double xj = x[j];
        const double t = clampd(xj - tau * g, a.lb[j], a.ub[j]);
        const double d = t - xj;
        local += d * d;
        if (kSingle)
            a.dx32[j] = float(d);
        else
            a.dx[j] = d;
        if (a.f.halpern) {
            xan[j] = t;
            xn[j] = lambda * a.anchor_x[j] + (1 - lambda) * ((1 + rho) * t - rho * xj);
        } else {
            xn[j] = t;
            const double v = xa[j];
            xan[j] = v + (t - v) * w;
        }
    };
    long long slot = blockIdx.x;
    if (a.part) for_rows_partials(at.rows, a.part, a.parts, update);
    else for_rows<G>(at, sg, y, lds, update, &slot);
    const double s = block_sum(local, lds);
    if (threadIdx.x == 0 && slot >= 0) store_shared(&a.partial[slot], s);
}

// Accept/reject and step-factor update (src/cpu.cpp:151-202).  Thread 0 only.
__device__ void decide(State *st, const Flags &fl, double dx2, double dy2, double coupling) {
    const double f = st->factor;
    const double tau = st->tau0 * f, sigma = st->sigma0 * f;
    const double energy = dx2 / tau + dy2 / sigma;
    const double limit = fabs(coupling) > 0 ? energy / (2 * fabs(coupling)) : INFINITY;
    const bool finite = isfinite(energy) && isfinite(coupling);
    const bool accept = finite && (!fl.adaptive || limit >= 1);
    if (fl.adaptive) {
        const double k = double(st->accepted + 2);
        double multiplier = mind((1 - pow(k, -.3)) * limit, 1 + pow(k, -.6));
        if (!finite || !isfinite(multiplier) || multiplier <= 0) multiplier = .5;
        st->factor = clampd(f * multiplier, 1e-12, 1e12);
    }
    if (!accept) {
        st->rejected += 1;
        st->consecutive_rejects += 1;
        if (!fl.adaptive)
            st->stop = kNonfinite;
        else if (st->consecutive_rejects >= kMaxRejects)
            st->stop = kExhausted;
        return;
    }
    st->cur ^= 1;
    st->accepted += 1;
    st->chunk_accepted += 1;
    st->consecutive_rejects = 0;
    if (fl.halpern) {
        const double residual = sqrt(maxd(0., energy - 2 * coupling));
        if (st->epoch_steps == 0) st->initial_residual = residual;
        st->epoch_steps += 1;
        const long long half = st->accepted / 2;
        const bool restart = fl.residual_restarts && st->epoch_steps >= 10 &&
                             (residual <= .2 * st->initial_residual ||
                              (residual <= .8 * st->initial_residual && residual > st->previous_residual) ||
                              st->epoch_steps >= (half > 10 ? half : 10));
        st->restart_requested = restart ? 1 : 0;
        st->previous_residual = residual;
        if (restart) {
            st->stop = kRestart;
            return;
        }
    } else {
        st->average_mass += f;
    }
    if (st->chunk_accepted >= st->chunk_target) st->stop = kDone;
}

// T: arithmetic of A dx (as for primal_kernel); in the FP32 phase the new y
// is also stored as FP32 for the next A^T y.
template <int G, class O, class V, class T = double>
__global__ __launch_bounds__(kBlock) void dual_kernel(DevCsr<O, V> A, Segments sg, DualArgs a) {
    constexpr bool kSingle = sizeof(T) == 4;
    State *st = a.st;
    if (st->stop != kRun) return;
    const int cur = st->cur;
    const double *y = a.y[cur];
    double *yn = a.y[cur ^ 1];
    float *yn32 = a.y32[cur ^ 1];
    const double *ya = a.ya[cur];
    double *yan = a.ya[cur ^ 1];
    const double *ax = a.ax[cur];
    double *axn = a.ax[cur ^ 1];
    const double f = st->factor;
    const double sigma = st->sigma0 * f;
    const double w = f / (st->average_mass + f);
    const double lambda = 1 / (double(st->epoch_steps) + 2);
    const double rho = a.f.reflection;
    double sq = 0, cross = 0;
    __shared__ double lds[4];
    __shared__ int last;
    auto update = [&](long long i, double delta, double) {
        const double axi = ax[i];
        const double bar = axi + 2 * delta;
        const double yi = y[i];
        const double z = yi + sigma * bar;
        const double t = maxd(0., z - sigma * a.ru[i]) + mind(0., z - sigma * a.rl[i]);
        const double dy = t - yi;
        sq += (!isfinite(bar) || !isfinite(t)) ? INFINITY : dy * dy;
        cross += dy * delta;
        const double next_ax = axi + delta;
        if (a.f.halpern) {
            yan[i] = t;
            const double next = lambda * a.anchor_y[i] + (1 - lambda) * ((1 + rho) * t - rho * yi);
            yn[i] = next;
            if (kSingle) yn32[i] = float(next);
            axn[i] = lambda * a.anchor_ax[i] + (1 - lambda) * ((1 + rho) * next_ax - rho * axi);
        } else {
            yn[i] = t;
            if (kSingle) yn32[i] = float(t);
            const double v = ya[i];
            yan[i] = v + (t - v) * w;
            axn[i] = next_ax;
        }
    };
    long long slot = blockIdx.x;
    if (a.part) for_rows_partials(A.rows, a.part, a.parts, update);
    else for_rows<G>(A, sg, gathered<T>(a.dx, a.dx32), lds, update, &slot);
    const double s1 = block_sum(sq, lds);
    const double s2 = block_sum(cross, lds);
    if (threadIdx.x == 0 && slot >= 0) {
        store_shared(&a.partial[2 * slot], s1);
        store_shared(&a.partial[2 * slot + 1], s2);
    }
    // Large grids hand the decision to decide_kernel: one ticket per block on
    // a single address serialises at the memory side (tens of ns each).
    if (!a.inline_decide) return;
    if (threadIdx.x == 0) {
        const unsigned ticket = __hip_atomic_fetch_add(&st->ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = ticket == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    // Every thread issues its share of the partial loads together (they are
    // memory-side round trips), then adds them in a fixed order.
    const double p = block_sum(gather_sum(a.primal_partial, 1, 0, a.primal_blocks), lds);
    const double d = block_sum(gather_sum(a.partial, 2, 0, (int)gridDim.x), lds);
    const double c = block_sum(gather_sum(a.partial, 2, 1, (int)gridDim.x), lds);
    if (threadIdx.x == 0) {
        decide(st, a.f, p, d, c);
        __hip_atomic_store(&st->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The accept/reject decision of one attempt after a dual_kernel with
// inline_decide == 0 (same partials, same fixed order).
__global__ __launch_bounds__(kBlock) void decide_kernel(DualArgs a, int dual_blocks) {
    State *st = a.st;
    if (st->stop != kRun) return;
    __shared__ double lds[4];
    const double p = block_sum(gather_sum(a.primal_partial, 1, 0, a.primal_blocks), lds);
    const double d = block_sum(gather_sum(a.partial, 2, 0, dual_blocks), lds);
    const double c = block_sum(gather_sum(a.partial, 2, 1, dual_blocks), lds);
    if (threadIdx.x == 0) decide(st, a.f, p, d, c);
}


template <int G, class O, class V>
__global__ __launch_bounds__(kBlock) void spmv_kernel(DevCsr<O, V> A, Segments sg, const double *x, double *out) {
    __shared__ double lds[4];
    for_rows<G>(A, sg, x, lds, [&](long long i, double acc, double) { out[i] = acc; });
}

// out = A x and out_abs = |A| |x| (the rounding scale of each entry of out).
template <int G, class O>
__global__ __launch_bounds__(kBlock) void spmv_abs_kernel(DevCsr<O, double> A, Segments sg, const double *x,
                                                          double *out, double *out_abs) {
    __shared__ double lds[4];
    for_rows<G, true>(A, sg, x, lds, [&](long long i, double acc, double acca) {
        out[i] = acc;
        out_abs[i] = acca;
    });
}

__global__ void begin_chunk_kernel(State *st, long long target, double tau0, double sigma0) {
    st->chunk_target = target;
    st->chunk_accepted = 0;
    st->stop = target > 0 ? kRun : kDone;
    st->tau0 = tau0;
    st->sigma0 = sigma0;
    st->consecutive_rejects = 0;
    st->ticket = 0;
}

// Restart bookkeeping (src/cpu.cpp:232-247): averages restart from zero mass,
// Halpern anchors and residual history are cleared.
__global__ void restart_state_kernel(State *st, int halpern) {
    st->average_mass = 0;
    if (halpern) {
        st->epoch_steps = 0;
        st->initial_residual = INFINITY;
        st->previous_residual = INFINITY;
        st->restart_requested = 0;
    }
}

__global__ void to_float_kernel(const double *src, float *dst, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = float(src[i]);
}

__global__ void fill_kernel(double *p, long long n, double v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void copy_kernel(double *dst, const double *src, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// --------------------------------------------------------- device checks --
//
// Original-space evaluation of a candidate with the formulas of the
// reference's verifier (src/verify.cpp:103-240, without the outward-rounded
// bound, which stays on the host), the Farkas pre-filter of the candidate's
// dual as a ray, and the epoch displacement.  Long sums are double-double,
// every reduction has a fixed order.

struct DD {
    double hi, lo;
};
__device__ __forceinline__ void dd_add(DD &a, double b) {
    const double s = a.hi + b;
    const double bb = s - a.hi;
    const double e = (a.hi - (s - bb)) + (b - bb);
    a.hi = s;
    a.lo += e;
}
__device__ __forceinline__ DD dd_merge(DD a, DD b) {
    const double s = a.hi + b.hi;
    const double bb = s - a.hi;
    double e = (a.hi - (s - bb)) + (b.hi - bb);
    e += a.lo + b.lo;
    const double h = s + e;
    return DD{h, e - (h - s)};
}
__device__ __forceinline__ double dd_value(DD a) { return a.hi + a.lo; }

constexpr double kEps = 2.220446049250313e-16;

// Per-thread sums, maxima and double-double sums.
template <int NS, int NM, int ND>
struct Terms {
    static constexpr int kWords = NS + NM + 2 * ND;
    double s[NS > 0 ? NS : 1];
    double mx[NM > 0 ? NM : 1];
    DD d[ND > 0 ? ND : 1];

    __device__ void merge(const Terms &o) {
#pragma unroll
        for (int k = 0; k < NS; ++k) s[k] += o.s[k];
#pragma unroll
        for (int k = 0; k < NM; ++k) mx[k] = maxd(mx[k], o.mx[k]);
#pragma unroll
        for (int k = 0; k < ND; ++k) d[k] = dd_merge(d[k], o.d[k]);
    }
    __device__ void store(double *p) const {
#pragma unroll
        for (int k = 0; k < NS; ++k) store_shared(p + k, s[k]);
#pragma unroll
        for (int k = 0; k < NM; ++k) store_shared(p + NS + k, mx[k]);
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            store_shared(p + NS + NM + 2 * k, d[k].hi);
            store_shared(p + NS + NM + 2 * k + 1, d[k].lo);
        }
    }
    __device__ void load(const double *p) {
#pragma unroll
        for (int k = 0; k < NS; ++k) s[k] = load_shared(p + k);
#pragma unroll
        for (int k = 0; k < NM; ++k) mx[k] = load_shared(p + NS + k);
#pragma unroll
        for (int k = 0; k < ND; ++k) d[k] = DD{load_shared(p + NS + NM + 2 * k), load_shared(p + NS + NM + 2 * k + 1)};
    }
};

// Reduces over the 256-thread block; the result is valid in thread 0.
template <int NS, int NM, int ND>
__device__ void reduce_block(Terms<NS, NM, ND> &t, double *lds) {
    constexpr int W = Terms<NS, NM, ND>::kWords;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < NS; ++k) t.s[k] += __shfl_xor(t.s[k], off, 64);
#pragma unroll
        for (int k = 0; k < NM; ++k) t.mx[k] = maxd(t.mx[k], __shfl_xor(t.mx[k], off, 64));
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            const DD o{__shfl_xor(t.d[k].hi, off, 64), __shfl_xor(t.d[k].lo, off, 64)};
            t.d[k] = dd_merge(t.d[k], o);
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        double *p = lds + wave * W;
#pragma unroll
        for (int k = 0; k < NS; ++k) p[k] = t.s[k];
#pragma unroll
        for (int k = 0; k < NM; ++k) p[NS + k] = t.mx[k];
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            p[NS + NM + 2 * k] = t.d[k].hi;
            p[NS + NM + 2 * k + 1] = t.d[k].lo;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w) {
            const double *p = lds + w * W;
            Terms<NS, NM, ND> o;
#pragma unroll
            for (int k = 0; k < NS; ++k) o.s[k] = p[k];
#pragma unroll
            for (int k = 0; k < NM; ++k) o.mx[k] = p[NS + k];
#pragma unroll
            for (int k = 0; k < ND; ++k) o.d[k] = DD{p[NS + NM + 2 * k], p[NS + NM + 2 * k + 1]};
            t.merge(o);
        }
    }
    __syncthreads();
}

// Grid-wide fixed-order reduction: every block stores its terms; the last
// block to finish merges them in block order.  Returns true in thread 0 of
// that block, with the grid total in t.
template <int NS, int NM, int ND>
__device__ bool reduce_grid(Terms<NS, NM, ND> &t, double *partial, unsigned *ticket, double *lds) {
    constexpr int W = Terms<NS, NM, ND>::kWords;
    __shared__ int last;
    reduce_block(t, lds);
    if (threadIdx.x == 0) {
        t.store(partial + (size_t)blockIdx.x * W);
        const unsigned k = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = k == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return false;
    Terms<NS, NM, ND> acc{};
    for (int b = threadIdx.x; b < (int)gridDim.x; b += kBlock) {
        Terms<NS, NM, ND> o;
        o.load(partial + (size_t)b * W);
        acc.merge(o);
    }
    reduce_block(acc, lds);
    if (threadIdx.x == 0) {
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        t = acc;
    }
    return threadIdx.x == 0;
}

// restore_x (src/preprocess.cpp:121-126) through the inverse column map.
__global__ void restore_x_kernel(const double *x, const int *colinv, const double *dc, const double *fixed,
                                 double *ox, long long n0) {
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n0; j += (long long)gridDim.x * blockDim.x) {
        const int k = colinv[j];
        ox[j] = k >= 0 ? x[k] * dc[k] : fixed[j];
    }
}

// restore_y (src/preprocess.cpp:127-132): each original row receives at most
// one kept row as its upper and one as its lower representative.
__global__ void restore_y_kernel(const double *y, const int *upinv, const int *loinv, const double *dr, double sobj,
                                 double *oy, long long m0) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m0; i += (long long)gridDim.x * blockDim.x) {
        double v = 0.0;
        const int u = upinv[i], l = loinv[i];
        if (u >= 0 && y[u] >= 0) v += y[u] * dr[u] * sobj;
        if (l >= 0 && !(y[l] >= 0)) v += y[l] * dr[l] * sobj;
        oy[i] = v;
    }
}

__global__ void sub_kernel(const double *a, const double *b, double *out, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = a[i] - b[i];
}

// Upper bound of min over [l, u] of g * x (concave in g).
__device__ __forceinline__ double phi(double g, double l, double u) {
    if (g > 0) return isfinite(l) ? g * l : -INFINITY;
    if (g < 0) return isfinite(u) ? g * u : -INFINITY;
    return 0;
}

struct EvalArgs {
    long long m0, n0;
    const double *ox, *oy, *act, *aty, *atyabs;
    const double *c0, *lb0, *ub0, *rl0, *ru0;
    const int *collen;
    double offset0, cscale;
    double *partial;
    unsigned *ticket;
    npdhg_accuracy *out;
};

// s : objective, Farkas column bound sum, its |.| sum, Lipschitz slack, sum |y b|
// mx: nonfinite, Farkas column -inf, primal, primal_abs, dual, dual_abs,
//     y != 0 at an infinite row endpoint, infinite Lagrangian box argument
// d : support, complement, box_min, sum y b over rows
using EvalTerms = Terms<5, 8, 4>;

__global__ __launch_bounds__(kBlock) void eval_kernel(EvalArgs a) {
    EvalTerms r{};
    const long long stride = (long long)gridDim.x * kBlock;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < a.m0; i += stride) {
        const double ax = a.act[i], yi = a.oy[i], lo = a.rl0[i], hi = a.ru0[i];
        if (!isfinite(ax) || !isfinite(yi)) r.mx[0] = 1;
        const double v = maxd(maxd(0., lo - ax), ax - hi);
        if (isfinite(lo)) r.mx[2] = maxd(r.mx[2], maxd(0., lo - ax) / (1 + fabs(lo)));
        if (isfinite(hi)) r.mx[2] = maxd(r.mx[2], maxd(0., ax - hi) / (1 + fabs(hi)));
        r.mx[3] = maxd(r.mx[3], v);
        if (yi != 0) {
            const double endpoint = yi > 0 ? hi : lo;
            if (isfinite(endpoint)) {
                dd_add(r.d[0], yi * endpoint);
                dd_add(r.d[1], fabs(yi * (endpoint - ax)));
                dd_add(r.d[3], yi * endpoint);
                r.s[4] += fabs(yi * endpoint);
            } else {
                r.mx[6] = 1;
                r.mx[4] = maxd(r.mx[4], fabs(yi));
                r.mx[5] = maxd(r.mx[5], fabs(yi));
            }
        }
    }
    for (long long j = (long long)blockIdx.x * kBlock + threadIdx.x; j < a.n0; j += stride) {
        const double xj = a.ox[j], l = a.lb0[j], u = a.ub0[j], c = a.c0[j];
        if (!isfinite(xj)) r.mx[0] = 1;
        const double v = maxd(maxd(0., l - xj), xj - u);
        r.mx[2] = maxd(r.mx[2], v / (1 + fabs(xj)));
        r.mx[3] = maxd(r.mx[3], v);
        r.s[0] += c * xj;
        const double linear = c + a.aty[j];
        if (!isfinite(linear)) r.mx[0] = 1;
        const double d = clampd(linear, xj - u, xj - l);
        r.mx[5] = maxd(r.mx[5], fabs(d));
        r.mx[4] = maxd(r.mx[4], fabs(d) / a.cscale);
        if (linear > 0 && isfinite(l)) dd_add(r.d[1], fabs(linear * (xj - l)));
        if (linear < 0 && isfinite(u)) dd_add(r.d[1], fabs(linear * (u - xj)));
        const double arg = linear > 0 ? l : linear < 0 ? u : clampd(0., l, u);
        if (!isfinite(arg)) r.mx[7] = 1;
        else dd_add(r.d[2], linear * arg);
        // Farkas column bound of the dual as a ray, widened by the rounding of A^T y.
        const double g = a.aty[j], ga = a.atyabs[j];
        const double e = double(a.collen[j] + 4) * 2 * kEps * ga;
        double ubd = maxd(phi(g - e, l, u), phi(g + e, l, u));
        if (g - e <= 0 && g + e >= 0) ubd = maxd(ubd, 0.0);
        if (ubd == -INFINITY) r.mx[1] = 1;
        else {
            r.s[1] += ubd;
            r.s[2] += fabs(ubd);
            const double box = maxd(isfinite(l) ? fabs(l) : 0.0, isfinite(u) ? fabs(u) : 0.0);
            r.s[3] += ga * box;
        }
    }
    __shared__ double lds[(kBlock / 64) * EvalTerms::kWords];
    if (!reduce_grid(r, a.partial, a.ticket, lds)) return;
    npdhg_accuracy out{};
    const double support = dd_value(r.d[0]), complement = dd_value(r.d[1]), box_min = dd_value(r.d[2]);
    out.objective = a.offset0 + r.s[0];
    out.primal = r.mx[2];
    out.dual = r.mx[4];
    out.gap = complement / (1 + fabs(out.objective - a.offset0));
    if (r.mx[6] == 0 && r.mx[7] == 0) {
        const double lower = (a.offset0 - support) + box_min;
        const double guard = 64 * kEps * (1 + fabs(a.offset0) + fabs(support) + fabs(box_min));
        const double lb = lower - guard;
        out.gap = maxd(out.gap, fabs(out.objective - lb) / (1 + fabs(out.objective) + fabs(lb)));
    }
    out.kkt = maxd(maxd(out.primal, out.dual), out.gap);
    out.finite = r.mx[0] == 0 && isfinite(out.objective) && isfinite(out.kkt);
    if (!out.finite) out.kkt = INFINITY;
    const bool neg_inf = r.mx[1] != 0 || r.mx[6] != 0;
    const double value = r.s[1] - dd_value(r.d[3]);
    const double slack = double(a.n0 + a.m0 + 16) * 4 * kEps * (r.s[2] + r.s[4] + r.s[3]);
    out.farkas_possible = !(neg_inf || value + slack <= 0);
    *a.out = out;
}

// Squared displacement |x - ex|^2 and |y - ey|^2, products split exactly.
__global__ __launch_bounds__(kBlock) void displacement_kernel(const double *x, const double *ex, long long n,
                                                              const double *y, const double *ey, long long m,
                                                              double *partial, unsigned *ticket, double *out) {
    Terms<0, 0, 2> r{};
    const long long stride = (long long)gridDim.x * kBlock;
    for (long long j = (long long)blockIdx.x * kBlock + threadIdx.x; j < n; j += stride) {
        const double d = x[j] - ex[j], p = d * d;
        dd_add(r.d[0], p);
        r.d[0].lo += fma(d, d, -p);
    }
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < m; i += stride) {
        const double d = y[i] - ey[i], p = d * d;
        dd_add(r.d[1], p);
        r.d[1].lo += fma(d, d, -p);
    }
    __shared__ double lds[(kBlock / 64) * Terms<0, 0, 2>::kWords];
    if (!reduce_grid(r, partial, ticket, lds)) return;
    out[0] = dd_value(r.d[0]);
    out[1] = dd_value(r.d[1]);
}

// ------------------------------------------------------------------ host --

int pick_group(long long rows, long long nnz) {
    double avg = rows > 0 ? double(nnz) / double(rows) : 1;
    int g = 1;
    while (g < 64 && g < avg / 2) g *= 2;
    return g;
}

int grid_for(long long rows, int group) {
    long long rpb = kBlock / group;
    long long blocks = (rows + rpb - 1) / rpb;
    return int(std::max<long long>(1, std::min<long long>(blocks, kMaxBlocks)));
}

constexpr long long kSegment = 2048;  // nonzeros per long-row segment (8 per thread)

// Longest row a lane group of `group` lanes walks itself (32 rounds).
long long row_threshold(int group) { return std::max<long long>(32LL * group, 128); }

struct HostSegments {
    std::vector<long long> begin, end;
    std::vector<int> owner, row, first, count;
};

HostSegments make_segments(const int64_t *ptr, long long rows, long long threshold) {
    HostSegments h;
    for (long long i = 0; i < rows; ++i) {
        const long long len = ptr[i + 1] - ptr[i];
        if (len <= threshold) continue;
        const int slot = int(h.row.size());
        const long long pieces = (len + kSegment - 1) / kSegment;
        h.row.push_back(int(i));
        h.first.push_back(int(h.begin.size()));
        h.count.push_back(int(pieces));
        for (long long q = 0; q < pieces; ++q) {
            h.begin.push_back(ptr[i] + q * kSegment);
            h.end.push_back(std::min<long long>(ptr[i] + (q + 1) * kSegment, ptr[i + 1]));
            h.owner.push_back(slot);
        }
        if (h.begin.size() > size_t(1) << 30) throw std::runtime_error("Too many long-row segments");
    }
    return h;
}

// Gathered vectors up to this size stay in one XCD's L2 (4 MB) next to the
// streamed matrix; larger ones are split into chunks of at most this size.
constexpr long long kChunkBytes = 2LL << 20;
// Below this size a gathered vector is served well enough by the L2s and the
// memory-side cache that the partial planes of a chunked product cost more
// than they save (measured on 2.9-8 MB vectors of 3-6M-nonzero LPs).
constexpr long long kChunkMinVector = 12LL << 20;

int chunk_count(long long vector_length) {
    if (const char *e = std::getenv("NIRYUKTI_XCD_CHUNKS"))
        if (std::atoi(e) == 0) return 0;
    long long min_bytes = kChunkMinVector;
    if (const char *e = std::getenv("NIRYUKTI_XCD_CHUNK_MIN_MB")) min_bytes = std::atoll(e) << 20;
    if (vector_length * 8 < min_bytes) return 0;
    int p = 1;
    while (p < 8 && vector_length * 8 / p > kChunkBytes) p *= 2;
    return p > 1 ? p : 0;
}

// Mean number of distinct 128-byte lines of the gathered vector that a window
// of 8,192 consecutive rows touches (rows longer than `threshold` excluded:
// they run as segments), in bytes.  A matrix whose windows stay within a
// fraction of one XCD's L2 (block-diagonal structure, banded rows) gathers
// from L2 without column chunks, and chunking would only add the partial
// planes' traffic.
double gather_footprint(const int64_t *ptr, const int64_t *index, long long rows, long long cols,
                        long long threshold) {
    constexpr long long kWindow = 8192;
    std::vector<int> stamp(static_cast<size_t>((cols + 15) / 16 + 1), -1);
    double lines = 0;
    long long windows = 0;
    for (long long w0 = 0; w0 < rows; w0 += kWindow, ++windows) {
        const int id = int(w0 / kWindow);
        for (long long i = w0; i < std::min(rows, w0 + kWindow); ++i) {
            if (ptr[i + 1] - ptr[i] > threshold) continue;
            for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e) {
                int &st = stamp[size_t(index[e] / 16)];
                if (st != id) {
                    st = id;
                    lines += 1;
                }
            }
        }
    }
    return windows ? 128.0 * lines / double(windows) : 0.0;
}
constexpr double kLocalFootprint = 1 << 20;  // bytes per window below which rows gather from L2 unchunked

struct HostChunks {
    int P = 0;
    std::vector<int> ptr;  // P * (rows + 1)
    std::vector<int> idx;
    std::vector<double> val;
    long long base[8] = {};
};

// Column chunks of a CSR matrix for partial_kernel; rows longer than
// `threshold` (segment rows) are left out.  Chunk boundaries split the
// columns so that every chunk holds about the same number of nonzeros: the
// chunks run side by side on different XCDs with equal shares of the grid,
// so a chunk that held the dense columns (the linking rows of a
// block-angular LP, for A^T) would otherwise keep one XCD busy long after
// the others finish.
HostChunks make_chunks(const int64_t *ptr, const int64_t *index, const double *value, long long rows,
                       long long cols, long long threshold, int P) {
    HostChunks h;
    h.P = P;
    std::vector<long long> per_col(static_cast<size_t>(cols) + 1, 0);
    long long kept = 0;
    for (long long i = 0; i < rows; ++i) {
        if (ptr[i + 1] - ptr[i] > threshold) continue;
        for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e) per_col[size_t(index[e])]++;
        kept += ptr[i + 1] - ptr[i];
    }
    // chunk_of[j]: the first k whose share of nonzeros the running count of
    // columns < j has not yet filled
    std::vector<unsigned char> chunk_of(static_cast<size_t>(cols));
    {
        long long run = 0;
        int k = 0;
        for (long long j = 0; j < cols; ++j) {
            while (k < P - 1 && run >= (long long)((__int128)kept * (k + 1) / P)) ++k;
            chunk_of[size_t(j)] = (unsigned char)k;
            run += per_col[size_t(j)];
        }
    }
    std::vector<long long> count(size_t(P) * size_t(rows + 1), 0);
    for (long long i = 0; i < rows; ++i) {
        if (ptr[i + 1] - ptr[i] > threshold) continue;
        for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e)
            count[size_t(chunk_of[size_t(index[e])]) * (rows + 1) + i + 1]++;
    }
    long long total = 0;
    h.ptr.resize(count.size());
    for (int k = 0; k < P; ++k) {
        h.base[k] = total;
        long long run = 0;
        for (long long i = 0; i <= rows; ++i) {
            run += count[size_t(k) * (rows + 1) + i];
            if (run >= (1LL << 31)) throw std::runtime_error("Chunk too large for 32-bit offsets");
            h.ptr[size_t(k) * (rows + 1) + i] = int(run);
        }
        total += run;
    }
    h.idx.resize(total);
    h.val.resize(total);
    std::vector<long long> fill(P);
    for (int k = 0; k < P; ++k) fill[k] = h.base[k];
    for (long long i = 0; i < rows; ++i) {
        if (ptr[i + 1] - ptr[i] > threshold) continue;
        for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e) {
            const int k = chunk_of[size_t(index[e])];
            h.idx[fill[k]] = int(index[e]);
            h.val[fill[k]] = value[e];
            ++fill[k];
        }
    }
    return h;
}

int elementwise_grid(long long n) { return int(std::max<long long>(1, std::min<long long>((n + 255) / 256, 4096))); }

struct HostCsr {
    long long rows = 0, cols = 0;
    std::vector<int64_t> ptr, index;
    std::vector<double> value;
};

HostCsr transpose(const npdhg_lp &lp) {
    HostCsr t;
    t.rows = lp.cols;
    t.cols = lp.rows;
    const int64_t nnz = lp.ptr[lp.rows];
    t.ptr.assign(lp.cols + 1, 0);
    t.index.resize(nnz);
    t.value.resize(nnz);
    for (int64_t k = 0; k < nnz; ++k) t.ptr[lp.index[k] + 1]++;
    for (int64_t j = 0; j < lp.cols; ++j) t.ptr[j + 1] += t.ptr[j];
    std::vector<int64_t> pos(t.ptr.begin(), t.ptr.end() - 1);
    for (int64_t i = 0; i < lp.rows; ++i)
        for (int64_t k = lp.ptr[i]; k < lp.ptr[i + 1]; ++k) {
            int64_t p = pos[lp.index[k]]++;
            t.index[p] = i;
            t.value[p] = lp.value[k];
        }
    return t;
}

class Backend {
  public:
    // `at`, when given, is the explicit transpose of lp's matrix (computed by
    // the caller, e.g. on the device); otherwise it is built here.
    Backend(const npdhg_lp &lp, const double *x0, const double *y0, const npdhg_config &cfg,
            const npdhg_lp *at = nullptr) {
        n_ = lp.cols;
        m_ = lp.rows;
        if (n_ < 0 || m_ < 0) throw std::runtime_error("Negative LP dimension");
        nnz_ = lp.ptr[lp.rows];
        if (cfg.halpern && cfg.adaptive) throw std::runtime_error("Halpern requires a fixed PDHG operator");
        if (!std::isfinite(cfg.reflection) || cfg.reflection < 0 || cfg.reflection > 1)
            throw std::runtime_error("Invalid Halpern reflection");
        if (n_ >= (1LL << 31) || m_ >= (1LL << 31))
            throw std::runtime_error("Dimensions beyond 2^31 rows/columns are not supported on the device");
        flags_.adaptive = cfg.adaptive;
        flags_.halpern = cfg.halpern;
        flags_.residual_restarts = cfg.residual_restarts;
        flags_.reflection = cfg.reflection;
        graphs_ = cfg.graphs != 0;
        fp32_ = cfg.fp32_values != 0;
        fp32_arith_ = cfg.fp32_arithmetic != 0;
        if (fp32_arith_ && fp32_) throw std::runtime_error("FP32 arithmetic keeps FP64 master values: not with fp32_values");
        if (fp32_arith_ && cfg.halpern) throw std::runtime_error("Fixed Halpern operators require FP64 arithmetic");
        const bool fits32 = nnz_ < (1LL << 31);
        if (cfg.index_bits == 32 && !fits32) throw std::runtime_error("CSR offsets do not fit 32 bits");
        idx64_ = cfg.index_bits == 64 || !fits32;
        HIP_OK(hipGetDevice(&device_));
        HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        try {
            upload_matrix(lp.rows, lp.cols, lp.ptr, lp.index, lp.value, a_);
            HostCsr built;
            const int64_t *tp, *ti;
            const double *tv;
            if (at) {
                if (at->rows != n_ || at->cols != m_ || !at->ptr || at->ptr[n_] != nnz_ || (nnz_ && (!at->index || !at->value)))
                    throw std::runtime_error("Transpose does not match the LP");
                tp = at->ptr;
                ti = at->index;
                tv = at->value;
            } else {
                built = transpose(lp);
                tp = built.ptr.data();
                ti = built.index.data();
                tv = built.value.data();
            }
            upload_matrix(n_, m_, tp, ti, tv, at_);
            g_a_ = pick_group(m_, nnz_);
            g_at_ = pick_group(n_, nnz_);
            sg_a_ = upload_segments(make_segments(lp.ptr, m_, row_threshold(g_a_)), row_threshold(g_a_));
            sg_at_ = upload_segments(make_segments(tp, n_, row_threshold(g_at_)), row_threshold(g_at_));
            nb1_ = sg_at_.nseg + grid_for(n_, g_at_);
            nb2_ = sg_a_.nseg + grid_for(m_, g_a_);
            setup_chunks(lp.ptr, lp.index, lp.value, m_, n_, row_threshold(g_a_), ch_a_, ch_a_val_, g_ch_a_,
                         nbp_a_, part_a_, ch_a_nnz_);
            setup_chunks(tp, ti, tv, n_, m_, row_threshold(g_at_), ch_at_,
                         ch_at_val_, g_ch_at_, nbp_at_, part_at_, ch_at_nnz_);
            if (fp32_arith_) {
                // FP32 copies of the values for the FP32 arithmetic phase; the
                // FP64 values stay for the refinement phase and the rebases
                a32_val_ = narrow(static_cast<const double *>(a_.val), a_.nnz);
                at32_val_ = narrow(static_cast<const double *>(at_.val), at_.nnz);
                if (ch_a_.P) ch_a32_val_ = narrow(static_cast<const double *>(ch_a_val_), ch_a_nnz_);
                if (ch_at_.P) ch_at32_val_ = narrow(static_cast<const double *>(ch_at_val_), ch_at_nnz_);
            }
            c_ = upload(lp.c, n_);
            lb_ = upload(lp.lb, n_);
            ub_ = upload(lp.ub, n_);
            rl_ = upload(lp.rl, m_);
            ru_ = upload(lp.ru, m_);
            for (int h = 0; h < 2; ++h) {
                x_[h] = alloc_d(n_);
                xa_[h] = alloc_d(n_);
                y_[h] = alloc_d(m_);
                ya_[h] = alloc_d(m_);
                ax_[h] = alloc_d(m_);
            }
            dx_ = alloc_d(n_);
            if (fp32_arith_) {
                for (int h = 0; h < 2; ++h) y32_[h] = static_cast<float *>(alloc_bytes(size_t(std::max<long long>(m_, 1)) * 4));
                dx32_ = static_cast<float *>(alloc_bytes(size_t(std::max<long long>(n_, 1)) * 4));
                phase32_ = true;
            }
            if (flags_.halpern) {
                anchor_x_ = alloc_d(n_);
                anchor_y_ = alloc_d(m_);
                anchor_ax_ = alloc_d(m_);
            }
            part1_ = alloc_d(nb1_);
            part2_ = alloc_d(2LL * nb2_);
            // segment slots other than each long row's first are never written
            zero(part1_, nb1_);
            zero(part2_, 2LL * nb2_);
            st_ = static_cast<State *>(alloc_bytes(sizeof(State)));
            State init{};
            init.factor = 1;
            init.initial_residual = INFINITY;
            init.previous_residual = INFINITY;
            init.stop = kDone;
            HIP_OK(hipMemcpyAsync(st_, &init, sizeof init, hipMemcpyHostToDevice, stream_));
            host_ = init;
            for (int h = 0; h < 2; ++h) {
                zero(xa_[h], n_);
                zero(ya_[h], m_);
                zero(ax_[h], m_);
                zero(x_[h], n_);
                zero(y_[h], m_);
            }
            zero(dx_, n_);
            set_point(x0, y0);
            HIP_OK(hipStreamSynchronize(stream_));
        } catch (...) {
            release();
            throw;
        }
    }
    ~Backend() { release(); }

    int advance(int count, double tau0, double sigma0) {
        if (count <= 0) return 0;
        begin_chunk_kernel<<<1, 1, 0, stream_>>>(st_, count, tau0, sigma0);
        HIP_OK(hipGetLastError());
        const int attempts = count + kSpareAttempts;
        if (graphs_) {
            hipGraphExec_t exec = graph_for(attempts);
            HIP_OK(hipGraphLaunch(exec, stream_));
        } else {
            for (int i = 0; i < attempts; ++i) launch_attempt();
        }
        sync_state();
        // Every attempt either accepts (bounded by count) or rejects (bounded
        // by kMaxRejects in a row), so this loop ends; the round limit guards
        // against a decision that never happened.
        for (int round = 0; host_.stop == kRun; ++round) {
            if (round > count + kMaxRejects)
                throw std::runtime_error("PDHG chunk made no progress on the device");
            for (int i = 0; i < 8; ++i) launch_attempt();
            sync_state();
        }
        if (host_.stop == kExhausted) throw std::runtime_error("PDHG backtracking exhausted");
        if (host_.stop == kNonfinite) throw std::runtime_error("Nonfinite PDHG step");
        return int(host_.chunk_accepted);
    }

    void status(npdhg_status *s) {
        sync_state();
        s->accepted_total = host_.accepted;
        s->rejected_total = host_.rejected;
        s->restart_requested = host_.restart_requested;
        s->index_bits = idx64_ ? 64 : 32;
        s->factor = host_.factor;
        s->fp32_phase = phase32_ ? 1 : 0;
        s->pad = 0;
    }

    void candidates(double *x, double *y, double *xa, double *ya) {
        sync_state();
        const int c = host_.cur;
        download(x, x_[c], n_);
        download(y, y_[c], m_);
        download(xa, xa_[c], n_);
        download(ya, ya_[c], m_);
        spmv(false, x_[c], ax_[c]);  // rebase the activity recurrence (src/cpu.cpp:229-230)
        HIP_OK(hipStreamSynchronize(stream_));
    }

    void reset(const double *x, const double *y) {
        sync_state();
        const int c = host_.cur;
        if (n_) HIP_OK(hipMemcpyAsync(x_[c], x, n_ * sizeof(double), hipMemcpyHostToDevice, stream_));
        if (m_) HIP_OK(hipMemcpyAsync(y_[c], y, m_ * sizeof(double), hipMemcpyHostToDevice, stream_));
        restart_resident(c);
        HIP_OK(hipStreamSynchronize(stream_));
    }

    void reset_candidate(bool average) {
        sync_state();
        const int c = host_.cur;
        if (average) {
            copy(x_[c], xa_[c], n_);
            copy(y_[c], ya_[c], m_);
        }
        restart_resident(c);
        HIP_OK(hipStreamSynchronize(stream_));
    }

    void spmv_host(bool transpose, const double *x, double *y) {
        long long in = transpose ? m_ : n_, out = transpose ? n_ : m_;
        double *dx = alloc_d(in), *dy = alloc_d(out);
        if (in) HIP_OK(hipMemcpyAsync(dx, x, in * sizeof(double), hipMemcpyHostToDevice, stream_));
        spmv(transpose, dx, dy);
        download(y, dy, out);
        free_one(dx);
        free_one(dy);
    }

    int index_bits() const { return idx64_ ? 64 : 32; }

    void attach_original(const npdhg_original &o) {
        if (attached_) throw std::runtime_error("Original model already attached");
        if (o.rows < 0 || o.cols < 0 || o.rows >= (1LL << 31) || o.cols >= (1LL << 31))
            throw std::runtime_error("Original model dimensions out of range");
        n0_ = o.cols;
        m0_ = o.rows;
        std::vector<int> colinv(n0_, -1), upinv(m0_, -1), loinv(m0_, -1);
        for (long long j = 0; j < n_; ++j) {
            const int64_t t = o.cols_map[j];
            if (t < 0 || t >= n0_ || colinv[t] >= 0) throw std::runtime_error("Column map is not one-to-one");
            colinv[t] = int(j);
        }
        for (long long i = 0; i < m_; ++i) {
            const int64_t u = o.upper_rows[i], l = o.lower_rows[i];
            if (u < 0 || u >= m0_ || l < 0 || l >= m0_ || upinv[u] >= 0 || loinv[l] >= 0)
                throw std::runtime_error("Row maps are not one-to-one");
            upinv[u] = int(i);
            loinv[l] = int(i);
        }
        const long long nnz0 = o.ptr[o.rows];
        wide0_ = nnz0 >= (1LL << 31);
        npdhg_lp view{};
        view.rows = o.rows;
        view.cols = o.cols;
        view.ptr = o.ptr;
        view.index = o.index;
        view.value = o.value;
        upload_matrix(o.rows, o.cols, o.ptr, o.index, o.value, a0_, wide0_, false);
        HostCsr t = transpose(view);
        upload_matrix(t.rows, t.cols, t.ptr.data(), t.index.data(), t.value.data(), at0_, wide0_, false);
        g_a0_ = pick_group(m0_, nnz0);
        g_at0_ = pick_group(n0_, nnz0);
        sg_a0_ = upload_segments(make_segments(o.ptr, m0_, row_threshold(g_a0_)), row_threshold(g_a0_));
        sg_at0_ = upload_segments(make_segments(t.ptr.data(), n0_, row_threshold(g_at0_)), row_threshold(g_at0_));
        std::vector<int> collen(n0_);
        for (long long j = 0; j < n0_; ++j) collen[j] = int(t.ptr[j + 1] - t.ptr[j]);
        collen_ = upload_vector(collen);
        colinv_ = upload_vector(colinv);
        upinv_ = upload_vector(upinv);
        loinv_ = upload_vector(loinv);
        c0_ = upload(o.c, n0_);
        lb0_ = upload(o.lb, n0_);
        ub0_ = upload(o.ub, n0_);
        rl0_ = upload(o.rl, m0_);
        ru0_ = upload(o.ru, m0_);
        fixed_ = upload(o.fixed, n0_);
        dc_ = upload(o.column_scale, n_);
        dr_ = upload(o.row_scale, m_);
        sobj_ = o.objective_scale;
        offset0_ = o.offset;
        cscale_ = 1;
        for (long long j = 0; j < n0_; ++j) cscale_ = std::max(cscale_, 1 + std::abs(o.c[j]));
        for (int k = 0; k < 2; ++k) {
            ox_[k] = alloc_d(n0_);
            oy_[k] = alloc_d(m0_);
        }
        bx_ = alloc_d(n0_);
        by_ = alloc_d(m0_);
        act_ = alloc_d(m0_);
        aty_ = alloc_d(n0_);
        atyabs_ = alloc_d(n0_);
        oyd_ = alloc_d(m0_);
        dtmp_ = alloc_d(m_);
        ex_ = alloc_d(n_);
        ey_ = alloc_d(m_);
        eval_blocks_ = int(std::max<long long>(1, std::min<long long>((std::max(n0_, m0_) + kBlock - 1) / kBlock, 1024)));
        eval_partial_ = alloc_d((long long)eval_blocks_ * EvalTerms::kWords);
        eval_ticket_ = static_cast<unsigned *>(alloc_bytes(2 * sizeof(unsigned)));
        HIP_OK(hipMemsetAsync(eval_ticket_, 0, 2 * sizeof(unsigned), stream_));
        eval_out_ = static_cast<npdhg_accuracy *>(alloc_bytes(3 * sizeof(npdhg_accuracy)));
        disp_out_ = alloc_d(2);
        sync_state();
        copy(ex_, x_[host_.cur], n_);
        copy(ey_, y_[host_.cur], m_);
        HIP_OK(hipStreamSynchronize(stream_));
        attached_ = true;
    }

    void evaluate(npdhg_accuracy *current, npdhg_accuracy *average) {
        need_original();
        sync_state();
        const int c = host_.cur;
        eval_point(x_[c], y_[c], ox_[0], oy_[0], eval_out_);
        eval_point(xa_[c], ya_[c], ox_[1], oy_[1], eval_out_ + 1);
        spmv(false, x_[c], ax_[c]);  // rebase the activity recurrence, as candidates() does
        npdhg_accuracy out[2];
        HIP_OK(hipMemcpyAsync(out, eval_out_, sizeof out, hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
        if (current) *current = out[0];
        if (average) *average = out[1];
        evaluated_ = true;
    }

    void keep_best(bool average) {
        need_evaluated();
        copy(bx_, ox_[average ? 1 : 0], n0_);
        copy(by_, oy_[average ? 1 : 0], m0_);
        HIP_OK(hipStreamSynchronize(stream_));
        best_ = true;
    }

    void original_point(int which, double *x0, double *y0) {
        need_evaluated();
        if (which < 0 || which > 2) throw std::runtime_error("Point must be 0 (current), 1 (average) or 2 (best)");
        if (which == 2 && !best_) throw std::runtime_error("No best point kept yet");
        const double *px = which == 2 ? bx_ : ox_[which], *py = which == 2 ? by_ : oy_[which];
        if (x0) download(x0, px, n0_);
        if (y0) download(y0, py, m0_);
    }

    void farkas_displacement(bool average, int *possible, double *ray) {
        need_evaluated();
        sync_state();
        const int c = host_.cur;
        const double *y = average ? ya_[c] : y_[c];
        if (m_) {
            sub_kernel<<<elementwise_grid(m_), 256, 0, stream_>>>(y, ey_, dtmp_, m_);
            HIP_OK(hipGetLastError());
        }
        if (m0_) {
            restore_y_kernel<<<elementwise_grid(m0_), 256, 0, stream_>>>(dtmp_, upinv_, loinv_, dr_, sobj_, oyd_, m0_);
            HIP_OK(hipGetLastError());
        }
        spmv0(true, oyd_, aty_, atyabs_);
        launch_eval(ox_[0], oyd_, eval_out_ + 2);
        npdhg_accuracy out{};
        HIP_OK(hipMemcpyAsync(&out, eval_out_ + 2, sizeof out, hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
        *possible = out.farkas_possible;
        if (out.farkas_possible && ray) download(ray, oyd_, m0_);
    }

    void epoch_displacement(bool average, double *dx2, double *dy2) {
        need_original();
        sync_state();
        const int c = host_.cur;
        displacement_kernel<<<eval_blocks_, kBlock, 0, stream_>>>(average ? xa_[c] : x_[c], ex_, n_,
                                                                  average ? ya_[c] : y_[c], ey_, m_, eval_partial_,
                                                                  eval_ticket_ + 1, disp_out_);
        HIP_OK(hipGetLastError());
        double out[2];
        HIP_OK(hipMemcpyAsync(out, disp_out_, sizeof out, hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
        *dx2 = out[0];
        *dy2 = out[1];
    }

    // Resumable state (IterationBackend::snapshot / restore,
    // src/cpu.cpp:49-89, src/cuda.cu:739-770), in doubles:
    //   x (n) | y (m) | xa (n) | ya (m) | ax (m) | anchor x (n) | anchor y (m) |
    //   anchor ax (m) | factor, average mass, initial residual, previous
    //   residual, accepted, rejected, epoch steps, restart requested |
    // and with device checks attached:
    //   epoch x (n) | epoch y (m) | best x (n0) | best y (m0) | best kept (0/1).
    // The anchors are zero unless the backend runs Halpern iterations.
    static constexpr int kSnapshotScalars = 8;
    long long snapshot_size() const {
        long long s = 3 * n_ + 5 * m_ + kSnapshotScalars;
        if (attached_) s += n_ + m_ + n0_ + m0_ + 1;
        return s;
    }

    void snapshot(double *out) {
        sync_state();
        const int c = host_.cur;
        double *at = out;
        auto take = [&](const double *dev, long long count) {
            if (dev)
                download(at, dev, count);
            else
                std::fill(at, at + count, 0.0);
            at += count;
        };
        take(x_[c], n_);
        take(y_[c], m_);
        take(xa_[c], n_);
        take(ya_[c], m_);
        take(ax_[c], m_);
        take(flags_.halpern ? anchor_x_ : nullptr, n_);
        take(flags_.halpern ? anchor_y_ : nullptr, m_);
        take(flags_.halpern ? anchor_ax_ : nullptr, m_);
        const double scalars[kSnapshotScalars] = {host_.factor,           host_.average_mass,
                                                  host_.initial_residual, host_.previous_residual,
                                                  double(host_.accepted), double(host_.rejected),
                                                  double(host_.epoch_steps), double(host_.restart_requested)};
        std::copy(scalars, scalars + kSnapshotScalars, at);
        at += kSnapshotScalars;
        if (attached_) {
            take(ex_, n_);
            take(ey_, m_);
            take(bx_, n0_);
            take(by_, m0_);
            *at++ = best_ ? 1.0 : 0.0;
        }
    }

    void restore(const double *in) {
        const long long total = snapshot_size();
        for (long long k = 0; k < total; ++k)
            if (std::isnan(in[k])) throw std::runtime_error("NaN in backend checkpoint state");
        const double *q = in + 3 * n_ + 5 * m_;
        auto counter = [](double v) { return v >= 0 && v < 9.0e15 && v == std::floor(v); };
        if (!(q[0] > 0) || !std::isfinite(q[0]) || !(q[1] >= 0) || !std::isfinite(q[1]) || !counter(q[4]) ||
            !counter(q[5]) || !counter(q[6]) || (q[7] != 0 && q[7] != 1))
            throw std::runtime_error("Invalid backend checkpoint counters");
        sync_state();
        const double *at = in;
        auto put = [&](double *dev, long long count) {
            if (dev && count) HIP_OK(hipMemcpyAsync(dev, at, size_t(count) * sizeof(double), hipMemcpyHostToDevice, stream_));
            at += count;
        };
        put(x_[0], n_);
        put(y_[0], m_);
        put(xa_[0], n_);
        put(ya_[0], m_);
        put(ax_[0], m_);
        put(flags_.halpern ? anchor_x_ : nullptr, n_);
        put(flags_.halpern ? anchor_y_ : nullptr, m_);
        put(flags_.halpern ? anchor_ax_ : nullptr, m_);
        at += kSnapshotScalars;
        State s = host_;
        s.factor = q[0];
        s.average_mass = q[1];
        s.initial_residual = q[2];
        s.previous_residual = q[3];
        s.accepted = (long long)q[4];
        s.rejected = (long long)q[5];
        s.epoch_steps = (long long)q[6];
        s.restart_requested = q[7] != 0;
        s.cur = 0;
        s.stop = kDone;
        s.chunk_accepted = s.chunk_target = 0;
        s.consecutive_rejects = 0;
        s.ticket = 0;
        HIP_OK(hipMemcpyAsync(st_, &s, sizeof s, hipMemcpyHostToDevice, stream_));
        refresh_y32(0);
        if (attached_) {
            put(ex_, n_);
            put(ey_, m_);
            put(bx_, n0_);
            put(by_, m0_);
            best_ = *at++ != 0;
            evaluated_ = false;
        }
        HIP_OK(hipStreamSynchronize(stream_));
        host_ = s;
    }

    void set_epoch(bool average) {
        need_original();
        sync_state();
        const int c = host_.cur;
        copy(ex_, average ? xa_[c] : x_[c], n_);
        copy(ey_, average ? ya_[c] : y_[c], m_);
        HIP_OK(hipStreamSynchronize(stream_));
    }

  private:
    struct DeviceMatrix {
        long long rows = 0, cols = 0, nnz = 0;
        void *ptr = nullptr;
        int *idx = nullptr;
        void *val = nullptr;
    };

    void *alloc_bytes(size_t bytes) {
        void *p = nullptr;
        HIP_OK(hipMalloc(&p, std::max<size_t>(bytes, 256)));
        allocations_.push_back(p);
        return p;
    }
    // Host-to-device copy ordered on this backend's stream, complete on
    // return (the source may be a temporary).  Nothing here touches the
    // legacy default stream, which other threads' graph captures forbid.
    void copy_in(void *dst, const void *src, size_t bytes) {
        HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
    }
    double *alloc_d(long long n) { return static_cast<double *>(alloc_bytes(size_t(n) * sizeof(double))); }
    void free_one(void *p) {
        auto it = std::find(allocations_.begin(), allocations_.end(), p);
        if (it != allocations_.end()) {
            HIP_OK(hipStreamSynchronize(stream_));
            hipFree(p);
            allocations_.erase(it);
        }
    }
    double *upload(const double *h, long long n) {
        double *d = alloc_d(n);
        if (n) HIP_OK(hipMemcpyAsync(d, h, size_t(n) * sizeof(double), hipMemcpyHostToDevice, stream_));
        return d;
    }
    void download(double *h, const double *d, long long n) {
        if (n) HIP_OK(hipMemcpyAsync(h, d, size_t(n) * sizeof(double), hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
    }
    void zero(double *p, long long n) {
        if (!n) return;
        fill_kernel<<<elementwise_grid(n), 256, 0, stream_>>>(p, n, 0.0);
        HIP_OK(hipGetLastError());
    }
    void copy(double *dst, const double *src, long long n) {
        if (!n) return;
        copy_kernel<<<elementwise_grid(n), 256, 0, stream_>>>(dst, src, n);
        HIP_OK(hipGetLastError());
    }

    void upload_matrix(long long rows, long long cols, const int64_t *ptr, const int64_t *index, const double *value,
                       DeviceMatrix &d) {
        upload_matrix(rows, cols, ptr, index, value, d, idx64_, fp32_);
    }
    void upload_matrix(long long rows, long long cols, const int64_t *ptr, const int64_t *index, const double *value,
                       DeviceMatrix &d, bool wide, bool single) {
        d.rows = rows;
        d.cols = cols;
        d.nnz = ptr[rows];
        for (long long k = 0; k < d.nnz; ++k)
            if (index[k] < 0 || index[k] >= cols) throw std::runtime_error("CSR column index out of range");
        for (long long i = 0; i < rows; ++i)
            if (ptr[i] > ptr[i + 1] || ptr[i] < 0) throw std::runtime_error("CSR offsets not monotone");
        if (ptr[0] != 0) throw std::runtime_error("CSR offsets must start at 0");
        if (wide) {
            d.ptr = alloc_bytes(size_t(rows + 1) * 8);
            HIP_OK(hipMemcpyAsync(d.ptr, ptr, size_t(rows + 1) * 8, hipMemcpyHostToDevice, stream_));
        } else {
            std::vector<int32_t> p32(ptr, ptr + rows + 1);
            d.ptr = alloc_bytes(size_t(rows + 1) * 4);
            copy_in(d.ptr, p32.data(), p32.size() * 4);
        }
        {
            std::vector<int32_t> i32(index, index + d.nnz);
            d.idx = static_cast<int *>(alloc_bytes(size_t(d.nnz) * 4));
            if (d.nnz) copy_in(d.idx, i32.data(), i32.size() * 4);
        }
        if (single) {
            std::vector<float> v32(value, value + d.nnz);
            d.val = alloc_bytes(size_t(d.nnz) * 4);
            if (d.nnz) copy_in(d.val, v32.data(), v32.size() * 4);
        } else {
            d.val = alloc_bytes(size_t(d.nnz) * 8);
            if (d.nnz) HIP_OK(hipMemcpyAsync(d.val, value, size_t(d.nnz) * 8, hipMemcpyHostToDevice, stream_));
        }
    }

    void setup_chunks(const int64_t *ptr, const int64_t *index, const double *value, long long rows, long long cols,
                      long long threshold, ChunkedCsr &c, void *&val, int &group, int &grid, double *&part,
                      long long &entries) {
        entries = 0;
        c = ChunkedCsr{};
        int P = chunk_count(cols);
        if (P && gather_footprint(ptr, index, rows, cols, threshold) < kLocalFootprint) P = 0;
        if (!P || rows == 0) return;
        HostChunks h = make_chunks(ptr, index, value, rows, cols, threshold, P);
        c.P = P;
        c.rows = rows;
        for (int k = 0; k < 8; ++k) c.base[k] = h.base[k];
        c.ptr = upload_vector(h.ptr);
        c.idx = upload_vector(h.idx);
        if (fp32_) val = upload_vector(std::vector<float>(h.val.begin(), h.val.end()));
        else val = upload_vector(h.val);
        entries = (long long)h.val.size();
        group = 1;
        const long long rpb = (long long)kBlock * kChunkRows;
        const long long per_chunk = std::min<long long>((rows + rpb - 1) / rpb, 4096 / P);
        const long long per_xcd = (per_chunk + (8 / P) - 1) / (8 / P);
        grid = int(8 * std::max<long long>(1, per_xcd));
        // P chunk planes plus the long-row plane, which stays zero except at long rows.
        part = alloc_d((long long)(P + 1) * rows);
        fill_kernel<<<elementwise_grid((P + 1) * rows), 256, 0, stream_>>>(part, (P + 1) * rows, 0.0);
        HIP_OK(hipGetLastError());
    }

    // An FP32 copy of an FP64 device array.
    float *narrow(const double *src, long long n) {
        float *dst = static_cast<float *>(alloc_bytes(size_t(std::max<long long>(n, 1)) * sizeof(float)));
        if (n) {
            to_float_kernel<<<elementwise_grid(n), 256, 0, stream_>>>(src, dst, n);
            HIP_OK(hipGetLastError());
        }
        return dst;
    }
    // FP32 copy of y_[c] (the vector A^T y gathers in the FP32 phase).
    void refresh_y32(int c) {
        if (!phase32_ || !m_) return;
        to_float_kernel<<<elementwise_grid(m_), 256, 0, stream_>>>(y_[c], y32_[c], m_);
        HIP_OK(hipGetLastError());
    }

    template <class T>
    T *upload_vector(const std::vector<T> &h) {
        T *d = static_cast<T *>(alloc_bytes(h.size() * sizeof(T)));
        if (!h.empty()) copy_in(d, h.data(), h.size() * sizeof(T));
        return d;
    }
    Segments upload_segments(const HostSegments &h, long long threshold) {
        Segments sg{};
        sg.nseg = int(h.begin.size());
        sg.threshold = threshold;
        sg.begin = upload_vector(h.begin);
        sg.end = upload_vector(h.end);
        sg.owner = upload_vector(h.owner);
        sg.row = upload_vector(h.row);
        sg.first = upload_vector(h.first);
        sg.count = upload_vector(h.count);
        sg.partial = alloc_d(2 * (long long)h.begin.size());
        sg.ticket = upload_vector(std::vector<unsigned>(h.row.size(), 0u));
        return sg;
    }

    template <class O, class V>
    DevCsr<O, V> view(const DeviceMatrix &d) const {
        return DevCsr<O, V>{static_cast<const O *>(d.ptr), d.idx, static_cast<const V *>(d.val), d.rows};
    }
    template <class O, class V>
    DevCsr<O, V> view_values(const DeviceMatrix &d, const void *values) const {
        return DevCsr<O, V>{static_cast<const O *>(d.ptr), d.idx, static_cast<const V *>(values), d.rows};
    }

    PrimalArgs primal_args() const {
        PrimalArgs p{};
        p.c = c_;
        p.lb = lb_;
        p.ub = ub_;
        for (int h = 0; h < 2; ++h) {
            p.x[h] = x_[h];
            p.xa[h] = xa_[h];
            p.y[h] = y_[h];
            p.y32[h] = y32_[h];
        }
        p.dx = dx_;
        p.dx32 = dx32_;
        p.anchor_x = anchor_x_;
        p.partial = part1_;
        p.part = ch_at_.P ? part_at_ : nullptr;
        p.parts = ch_at_.P;
        p.st = st_;
        p.f = flags_;
        return p;
    }
    DualArgs dual_args() const {
        DualArgs d{};
        d.rl = rl_;
        d.ru = ru_;
        for (int h = 0; h < 2; ++h) {
            d.y[h] = y_[h];
            d.ya[h] = ya_[h];
            d.ax[h] = ax_[h];
            d.y32[h] = y32_[h];
        }
        d.dx = dx_;
        d.dx32 = dx32_;
        d.anchor_y = anchor_y_;
        d.anchor_ax = anchor_ax_;
        d.partial = part2_;
        d.primal_partial = part1_;
        d.primal_blocks = nb1_;
        d.part = ch_a_.P ? part_a_ : nullptr;
        d.parts = ch_a_.P;
        d.inline_decide = nb2_ <= kInlineDecideBlocks ? 1 : 0;
        d.st = st_;
        d.f = flags_;
        return d;
    }

    // One attempt: A^T y with the primal step, A dx with the dual step and the
    // decision.  T = float is the FP32 arithmetic phase (FP32 value copies,
    // FP32 gathered vectors); the iterates stay FP64 either way.
    template <class O, class V, class T = double>
    void launch_attempt_t() {
        constexpr bool single = sizeof(T) == 4;
        auto at = single ? view_values<O, V>(at_, at32_val_) : view<O, V>(at_);
        auto a = single ? view_values<O, V>(a_, a32_val_) : view<O, V>(a_);
        const void *ch_at_val = single ? static_cast<const void *>(ch_at32_val_) : ch_at_val_;
        const void *ch_a_val = single ? static_cast<const void *>(ch_a32_val_) : ch_a_val_;
        auto pa = primal_args();
        auto da = dual_args();
        VecPair<T> yv, dxv;
        if constexpr (single) {
            yv = VecPair<T>{{y32_[0], y32_[1]}};
            dxv = VecPair<T>{{dx32_, dx32_}};
        } else {
            yv = VecPair<T>{{y_[0], y_[1]}};
            dxv = VecPair<T>{{dx_, dx_}};
        }
        if (ch_at_.P) launch_partial(ch_at_, ch_at_val, at, sg_at_, nbp_at_, yv, part_at_);
        switch (g_at_) {
#define PRIMAL_CASE(G) \
    case G: primal_kernel<G, O, V, T><<<nb1_, kBlock, 0, stream_>>>(at, sg_at_, pa); break;
            PRIMAL_CASE(1) PRIMAL_CASE(2) PRIMAL_CASE(4) PRIMAL_CASE(8) PRIMAL_CASE(16) PRIMAL_CASE(32) PRIMAL_CASE(64)
#undef PRIMAL_CASE
        }
        if (ch_a_.P) launch_partial(ch_a_, ch_a_val, a, sg_a_, nbp_a_, dxv, part_a_);
        switch (g_a_) {
#define DUAL_CASE(G) \
    case G: dual_kernel<G, O, V, T><<<nb2_, kBlock, 0, stream_>>>(a, sg_a_, da); break;
            DUAL_CASE(1) DUAL_CASE(2) DUAL_CASE(4) DUAL_CASE(8) DUAL_CASE(16) DUAL_CASE(32) DUAL_CASE(64)
#undef DUAL_CASE
        }
        if (!da.inline_decide) decide_kernel<<<1, kBlock, 0, stream_>>>(da, nb2_);
    }
    template <class O, class V, class T>
    void launch_partial(const ChunkedCsr &c, const void *val, const DevCsr<O, V> &m, const Segments &sg, int grid,
                        VecPair<T> vec, double *part) {
        partial_kernel<O, V, T><<<grid + sg.nseg, kBlock, 0, stream_>>>(c, static_cast<const V *>(val), m, sg, grid,
                                                                       st_, vec, part);
    }

    void launch_attempt() {
        if (idx64_) {
            if (phase32_) launch_attempt_t<int64_t, float, float>();
            else if (fp32_) launch_attempt_t<int64_t, float>();
            else launch_attempt_t<int64_t, double>();
        } else {
            if (phase32_) launch_attempt_t<int32_t, float, float>();
            else if (fp32_) launch_attempt_t<int32_t, float>();
            else launch_attempt_t<int32_t, double>();
        }
        HIP_OK(hipGetLastError());
    }

  public:
    // Switches the attempts between the FP32 arithmetic phase and FP64 (the
    // refinement phase); the iterates, averages and counters carry on.
    void set_fp32_phase(bool on) {
        if (on && !fp32_arith_) throw std::runtime_error("Backend was created without FP32 arithmetic");
        if (on == phase32_) return;
        sync_state();
        for (auto &kv : graphs_cache_) HIP_OK(hipGraphExecDestroy(kv.second));
        graphs_cache_.clear();
        phase32_ = on;
        refresh_y32(host_.cur);
        HIP_OK(hipStreamSynchronize(stream_));
    }
    bool fp32_phase() const { return phase32_; }

  private:

    template <class O, class V>
    void spmv_t(bool transpose, const double *x, double *out) {
        const DeviceMatrix &d = transpose ? at_ : a_;
        auto mat = view<O, V>(d);
        const int g = transpose ? g_at_ : g_a_;
        const Segments &sg = transpose ? sg_at_ : sg_a_;
        const int grid = sg.nseg + grid_for(d.rows, g);
        switch (g) {
#define SPMV_CASE(G) \
    case G: spmv_kernel<G, O, V><<<grid, kBlock, 0, stream_>>>(mat, sg, x, out); break;
            SPMV_CASE(1) SPMV_CASE(2) SPMV_CASE(4) SPMV_CASE(8) SPMV_CASE(16) SPMV_CASE(32) SPMV_CASE(64)
#undef SPMV_CASE
        }
    }
    void spmv(bool transpose, const double *x, double *out) {
        if ((transpose ? n_ : m_) == 0) return;
        if (idx64_) {
            if (fp32_) spmv_t<int64_t, float>(transpose, x, out); else spmv_t<int64_t, double>(transpose, x, out);
        } else {
            if (fp32_) spmv_t<int32_t, float>(transpose, x, out); else spmv_t<int32_t, double>(transpose, x, out);
        }
        HIP_OK(hipGetLastError());
    }

    hipGraphExec_t graph_for(int attempts) {
        auto it = graphs_cache_.find(attempts);
        if (it != graphs_cache_.end()) return it->second;
        hipGraph_t graph = nullptr;
        // Relaxed: other threads (other solves, the batched engine) keep
        // running their own streams while this one captures.
        HIP_OK(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed));
        try {
            for (int i = 0; i < attempts; ++i) launch_attempt();
        } catch (...) {
            hipGraph_t failed = nullptr;
            hipStreamEndCapture(stream_, &failed);
            if (failed) hipGraphDestroy(failed);
            throw;
        }
        HIP_OK(hipStreamEndCapture(stream_, &graph));
        hipGraphExec_t exec = nullptr;
        hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        HIP_OK(e);
        graphs_cache_[attempts] = exec;
        return exec;
    }

    void need_original() const {
        if (!attached_) throw std::runtime_error("Device checks need npdhg_attach_original first");
    }
    void need_evaluated() const {
        need_original();
        if (!evaluated_) throw std::runtime_error("No candidate evaluated yet");
    }

    template <class O>
    void spmv0_t(bool transpose, const double *x, double *out, double *out_abs) {
        const DeviceMatrix &d = transpose ? at0_ : a0_;
        auto mat = view<O, double>(d);
        const int g = transpose ? g_at0_ : g_a0_;
        const Segments &sg = transpose ? sg_at0_ : sg_a0_;
        const int grid = sg.nseg + grid_for(d.rows, g);
        if (out_abs) {
            switch (g) {
#define SPMV_ABS_CASE(G) \
    case G: spmv_abs_kernel<G, O><<<grid, kBlock, 0, stream_>>>(mat, sg, x, out, out_abs); break;
                SPMV_ABS_CASE(1) SPMV_ABS_CASE(2) SPMV_ABS_CASE(4) SPMV_ABS_CASE(8) SPMV_ABS_CASE(16)
                SPMV_ABS_CASE(32) SPMV_ABS_CASE(64)
#undef SPMV_ABS_CASE
            }
        } else {
            switch (g) {
#define SPMV0_CASE(G) \
    case G: spmv_kernel<G, O, double><<<grid, kBlock, 0, stream_>>>(mat, sg, x, out); break;
                SPMV0_CASE(1) SPMV0_CASE(2) SPMV0_CASE(4) SPMV0_CASE(8) SPMV0_CASE(16) SPMV0_CASE(32) SPMV0_CASE(64)
#undef SPMV0_CASE
            }
        }
    }
    // Products with the original model (FP64 values).
    void spmv0(bool transpose, const double *x, double *out, double *out_abs) {
        if ((transpose ? n0_ : m0_) == 0) return;
        if (wide0_) spmv0_t<int64_t>(transpose, x, out, out_abs);
        else spmv0_t<int32_t>(transpose, x, out, out_abs);
        HIP_OK(hipGetLastError());
    }

    void launch_eval(const double *ox, const double *oy, npdhg_accuracy *out) {
        EvalArgs e{};
        e.m0 = m0_;
        e.n0 = n0_;
        e.ox = ox;
        e.oy = oy;
        e.act = act_;
        e.aty = aty_;
        e.atyabs = atyabs_;
        e.c0 = c0_;
        e.lb0 = lb0_;
        e.ub0 = ub0_;
        e.rl0 = rl0_;
        e.ru0 = ru0_;
        e.collen = collen_;
        e.offset0 = offset0_;
        e.cscale = cscale_;
        e.partial = eval_partial_;
        e.ticket = eval_ticket_;
        e.out = out;
        eval_kernel<<<eval_blocks_, kBlock, 0, stream_>>>(e);
        HIP_OK(hipGetLastError());
    }

    // Restores (x, y) into (ox, oy), forms A0 ox and A0^T oy, evaluates.
    void eval_point(const double *x, const double *y, double *ox, double *oy, npdhg_accuracy *out) {
        if (n0_) {
            restore_x_kernel<<<elementwise_grid(n0_), 256, 0, stream_>>>(x, colinv_, dc_, fixed_, ox, n0_);
            HIP_OK(hipGetLastError());
        }
        if (m0_) {
            restore_y_kernel<<<elementwise_grid(m0_), 256, 0, stream_>>>(y, upinv_, loinv_, dr_, sobj_, oy, m0_);
            HIP_OK(hipGetLastError());
        }
        spmv0(false, ox, act_, nullptr);
        spmv0(true, oy, aty_, atyabs_);
        launch_eval(ox, oy, out);
    }

    void sync_state() {
        HIP_OK(hipMemcpyAsync(&host_, st_, sizeof(State), hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
    }

    void set_point(const double *x0, const double *y0) {
        if (n_) HIP_OK(hipMemcpyAsync(x_[0], x0, n_ * sizeof(double), hipMemcpyHostToDevice, stream_));
        if (m_) HIP_OK(hipMemcpyAsync(y_[0], y0, m_ * sizeof(double), hipMemcpyHostToDevice, stream_));
        spmv(false, x_[0], ax_[0]);
        refresh_y32(0);
        if (flags_.halpern) {
            copy(anchor_x_, x_[0], n_);
            copy(anchor_y_, y_[0], m_);
            copy(anchor_ax_, ax_[0], m_);
        }
    }

    // x_[c], y_[c] hold the restart point; rebuild activity, clear averages.
    void restart_resident(int c) {
        spmv(false, x_[c], ax_[c]);
        refresh_y32(c);
        zero(xa_[c], n_);
        zero(ya_[c], m_);
        if (flags_.halpern) {
            copy(anchor_x_, x_[c], n_);
            copy(anchor_y_, y_[c], m_);
            copy(anchor_ax_, ax_[c], m_);
        }
        restart_state_kernel<<<1, 1, 0, stream_>>>(st_, flags_.halpern);
        HIP_OK(hipGetLastError());
    }

    void release() {
        for (auto &kv : graphs_cache_) hipGraphExecDestroy(kv.second);
        graphs_cache_.clear();
        if (stream_) hipStreamSynchronize(stream_);
        for (void *p : allocations_) hipFree(p);
        allocations_.clear();
        if (stream_) hipStreamDestroy(stream_);
        stream_ = nullptr;
    }

    int device_ = 0;
    hipStream_t stream_ = nullptr;
    long long n_ = 0, m_ = 0, nnz_ = 0;
    bool idx64_ = false, fp32_ = false, graphs_ = true;
    // FP32 arithmetic phase: FP32 value copies, FP32 copies of y (by parity)
    // and dx; phase32_ says which arithmetic the attempts use now.
    bool fp32_arith_ = false, phase32_ = false;
    float *a32_val_ = nullptr, *at32_val_ = nullptr, *ch_a32_val_ = nullptr, *ch_at32_val_ = nullptr;
    float *y32_[2] = {}, *dx32_ = nullptr;
    long long ch_a_nnz_ = 0, ch_at_nnz_ = 0;
    int g_a_ = 1, g_at_ = 1, nb1_ = 1, nb2_ = 1;
    Flags flags_{};
    DeviceMatrix a_, at_;
    Segments sg_a_{}, sg_at_{};
    ChunkedCsr ch_a_{}, ch_at_{};  // XCD column chunks (P == 0: not used)
    void *ch_a_val_ = nullptr, *ch_at_val_ = nullptr;
    int g_ch_a_ = 1, g_ch_at_ = 1, nbp_a_ = 8, nbp_at_ = 8;
    double *part_a_ = nullptr, *part_at_ = nullptr;
    // Original model and device-check buffers (npdhg_attach_original).
    bool attached_ = false, evaluated_ = false, best_ = false, wide0_ = false;
    long long n0_ = 0, m0_ = 0;
    DeviceMatrix a0_, at0_;
    Segments sg_a0_{}, sg_at0_{};
    int g_a0_ = 1, g_at0_ = 1, eval_blocks_ = 1;
    int *colinv_ = nullptr, *upinv_ = nullptr, *loinv_ = nullptr, *collen_ = nullptr;
    double *c0_ = nullptr, *lb0_ = nullptr, *ub0_ = nullptr, *rl0_ = nullptr, *ru0_ = nullptr, *fixed_ = nullptr;
    double *dc_ = nullptr, *dr_ = nullptr;
    double sobj_ = 1, offset0_ = 0, cscale_ = 1;
    double *ox_[2] = {}, *oy_[2] = {}, *bx_ = nullptr, *by_ = nullptr, *act_ = nullptr, *aty_ = nullptr,
           *atyabs_ = nullptr, *oyd_ = nullptr, *dtmp_ = nullptr, *ex_ = nullptr, *ey_ = nullptr;
    double *eval_partial_ = nullptr, *disp_out_ = nullptr;
    unsigned *eval_ticket_ = nullptr;
    npdhg_accuracy *eval_out_ = nullptr;
    double *c_ = nullptr, *lb_ = nullptr, *ub_ = nullptr, *rl_ = nullptr, *ru_ = nullptr;
    double *x_[2] = {}, *xa_[2] = {}, *y_[2] = {}, *ya_[2] = {}, *ax_[2] = {};
    double *dx_ = nullptr, *anchor_x_ = nullptr, *anchor_y_ = nullptr, *anchor_ax_ = nullptr;
    double *part1_ = nullptr, *part2_ = nullptr;
    State *st_ = nullptr;
    State host_{};
    std::map<int, hipGraphExec_t> graphs_cache_;
    std::vector<void *> allocations_;
};

template <class F>
int guarded(F &&f) {
    try {
        g_error.clear();
        f();
        return 0;
    } catch (const std::exception &e) {
        g_error = e.what();
        return -1;
    } catch (...) {
        g_error = "unknown native error";
        return -1;
    }
}

}  // namespace

struct npdhg_backend {
    Backend impl;
    npdhg_backend(const npdhg_lp &lp, const double *x0, const double *y0, const npdhg_config &cfg,
                  const npdhg_lp *at = nullptr)
        : impl(lp, x0, y0, cfg, at) {}
};

extern "C" {

const char *npdhg_last_error(void) { return g_error.c_str(); }

int npdhg_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int npdhg_device_name(int device, char *name, int len) {
    return guarded([&] {
        hipDeviceProp_t prop;
        HIP_OK(hipGetDeviceProperties(&prop, device));
        std::snprintf(name, size_t(len), "%s (%s, %d CUs, %.0f GiB)", prop.name, prop.gcnArchName,
                      prop.multiProcessorCount, double(prop.totalGlobalMem) / (1024.0 * 1024 * 1024));
    });
}

int npdhg_set_fp32_phase(npdhg_backend *b, int on) {
    return guarded([&] {
        if (!b) throw std::runtime_error("Null argument");
        b->impl.set_fp32_phase(on != 0);
    });
}

int npdhg_snapshot_size(npdhg_backend *b, int64_t *doubles) {
    return guarded([&] {
        if (!b || !doubles) throw std::runtime_error("Null argument");
        *doubles = b->impl.snapshot_size();
    });
}

int npdhg_snapshot(npdhg_backend *b, double *state) {
    return guarded([&] {
        if (!b || !state) throw std::runtime_error("Null argument");
        b->impl.snapshot(state);
    });
}

int npdhg_restore(npdhg_backend *b, const double *state) {
    return guarded([&] {
        if (!b || !state) throw std::runtime_error("Null argument");
        b->impl.restore(state);
    });
}

int npdhg_create(const npdhg_lp *lp, const double *x0, const double *y0, const npdhg_config *cfg,
                 npdhg_backend **out) {
    return guarded([&] {
        if (!lp || !x0 || !y0 || !cfg || !out) throw std::runtime_error("Null argument");
        *out = new npdhg_backend(*lp, x0, y0, *cfg);
    });
}

int npdhg_create_transposed(const npdhg_lp *lp, const int64_t *t_ptr, const int64_t *t_index, const double *t_value,
                            const double *x0, const double *y0, const npdhg_config *cfg, npdhg_backend **out) {
    return guarded([&] {
        if (!lp || !t_ptr || !x0 || !y0 || !cfg || !out) throw std::runtime_error("Null argument");
        npdhg_lp at{};
        at.rows = lp->cols;
        at.cols = lp->rows;
        at.ptr = t_ptr;
        at.index = t_index;
        at.value = t_value;
        *out = new npdhg_backend(*lp, x0, y0, *cfg, &at);
    });
}

void npdhg_destroy(npdhg_backend *b) { delete b; }

int npdhg_advance(npdhg_backend *b, int count, double primal_step, double dual_step, int *accepted) {
    return guarded([&] {
        if (!b) throw std::runtime_error("Null backend");
        int a = b->impl.advance(count, primal_step, dual_step);
        if (accepted) *accepted = a;
    });
}

int npdhg_status_get(npdhg_backend *b, npdhg_status *s) {
    return guarded([&] { b->impl.status(s); });
}

int npdhg_candidates(npdhg_backend *b, double *x, double *y, double *xa, double *ya) {
    return guarded([&] { b->impl.candidates(x, y, xa, ya); });
}

int npdhg_reset(npdhg_backend *b, const double *x, const double *y) {
    return guarded([&] { b->impl.reset(x, y); });
}

int npdhg_reset_candidate(npdhg_backend *b, int average) {
    return guarded([&] { b->impl.reset_candidate(average != 0); });
}

int npdhg_attach_original(npdhg_backend *b, const npdhg_original *orig) {
    return guarded([&] {
        if (!b || !orig) throw std::runtime_error("Null argument");
        b->impl.attach_original(*orig);
    });
}

int npdhg_evaluate(npdhg_backend *b, npdhg_accuracy *current, npdhg_accuracy *average) {
    return guarded([&] {
        if (!b) throw std::runtime_error("Null backend");
        b->impl.evaluate(current, average);
    });
}

int npdhg_keep_best(npdhg_backend *b, int average) {
    return guarded([&] {
        if (!b) throw std::runtime_error("Null backend");
        b->impl.keep_best(average != 0);
    });
}

int npdhg_original_point(npdhg_backend *b, int which, double *x0, double *y0) {
    return guarded([&] {
        if (!b) throw std::runtime_error("Null backend");
        b->impl.original_point(which, x0, y0);
    });
}

int npdhg_farkas_displacement(npdhg_backend *b, int average, int *possible, double *ray) {
    return guarded([&] {
        if (!b || !possible) throw std::runtime_error("Null argument");
        b->impl.farkas_displacement(average != 0, possible, ray);
    });
}

int npdhg_epoch_displacement(npdhg_backend *b, int average, double *dx2, double *dy2) {
    return guarded([&] {
        if (!b || !dx2 || !dy2) throw std::runtime_error("Null argument");
        b->impl.epoch_displacement(average != 0, dx2, dy2);
    });
}

int npdhg_set_epoch(npdhg_backend *b, int average) {
    return guarded([&] {
        if (!b) throw std::runtime_error("Null backend");
        b->impl.set_epoch(average != 0);
    });
}

int npdhg_spmv(npdhg_backend *b, int transpose, const double *x, double *y) {
    return guarded([&] { b->impl.spmv_host(transpose != 0, x, y); });
}

}  // extern "C"
