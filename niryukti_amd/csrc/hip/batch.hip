// Batched and served restarted PDHG on MI355X (gfx950): one workgroup per LP.
//
// Every running LP is driven by one workgroup (256 or 1,024 threads by size
// class; the last LPs of a class share a persistent tail grid that takes
// them from a device-side queue), which runs the reference's continuous
// driver for that LP entirely on the device
// (src/solver.cpp:511-726): chunks of `check_every` accepted adaptive PDHG
// steps (src/cpu.cpp:96-222), activity rebasing, original-space evaluation of
// the current and averaged candidates (src/verify.cpp:103-240 formulae),
// best-candidate tracking, the restart rule and the displacement primal-weight
// update.  Certification stays on the host: a workgroup that finds a candidate
// with device kkt <= tol parks its LP, and the host re-evaluates it with the
// reference's FP64 verifier and outward-rounded Lagrangian bound before it is
// reported OPTIMAL; a park that fails certification resumes the LP exactly
// where the reference would continue (restart rule with the safe kkt).  The
// reference's Farkas test (src/solver.cpp:599-633) is preceded on the device by
// a conservative upper bound of the ray's Lagrangian value; only rays whose
// bound cannot be shown non-positive go to the host's exact check.
//
// Each LP is reduced by its own workgroup in a fixed order, so an LP's result
// is bit-identical whether it runs alone or among thousands of others.
//
// The host side admits LPs into a fixed HBM arena as room frees up
// (continuous batching between launches) and retires finished ones; its
// scheduling rounds are pipelined (a round is launched before the previous
// one's stopped LPs are certified).
#include <hip/hip_runtime.h>
#include <sys/resource.h>
#include <sys/syscall.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <exception>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine/lp.hpp"
#include "engine/parallel.hpp"
#include "niryukti_batch.h"

namespace {

thread_local std::string g_error;

#define HIP_OK(expr)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kThreads = 256;
// LDS-resident launch classes: capacity in doubles for y (m) plus dx (n);
// the last class runs the global-memory stream kernel.
constexpr int kClasses = 4;
constexpr int kClassDoubles[kClasses - 1] = {2048, 4096, 8000};
static_assert(kClassDoubles[kClasses - 2] <= 65536, "sliced entry indices are 16-bit");
// Upper bound of the LDS step kernels' static allocation (state, tile sums,
// flags; the fused tail kernel adds the check's reduction slots).
constexpr int kStaticLds = 8192;
constexpr int64_t kShortRow = 128;
static_assert(kShortRow <= 255, "LDS-class row lengths are stored in bytes");
// The two larger LDS classes (m + n above 2,048) run 1,024-thread workgroups
// with one row per thread: their LDS allows one or two workgroups per CU, and
// 16 waves per workgroup keep more loads in flight (measured: the largest
// class 8-10 % faster per period than 512 threads x 2 rows; the crowded
// phase 3 % shorter with the middle class moved too).
constexpr int kWideThreads = 1024;
// LDS classes with at most kTailLps LPs left (one per CU or fewer) are merged
// into one tail launch of kTailThreads-thread workgroups, one row per thread.
constexpr int kTailLps = 256;
constexpr int kTailThreads = 1024;
// A tail launch runs each LP for at least the requested check periods, then
// on until kTailBudgetMs have passed (at most kTailMaxPeriodsPerCheck times
// the requested periods), so that LPs of different speed do not wait for
// each other at every period.
constexpr int kTailBudgetMs = 32;
// Nice value of the prepare workers (measured on the streamed bench: the
// scheduling thread's share of the timed region outside its waits on the
// device fell from 0.45-0.68 s to 0.32-0.36 s, 7,278-7,428 -> 7,490-7,513
// LPs/s, with nice 5 or 10).
constexpr int kPrepareNice = 5;
// Rounds in which at most this many LPs stop certify them before returning;
// larger sets are certified on a host thread beside the next round.
constexpr int kInlineCertify = 32;
constexpr int kTailMaxPeriodsPerCheck = 1000;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxRejects = 40;
// Mixed precision (FP32 matrix values, FP64 products, sums and iterates): an
// LP moves to FP64 values when a candidate reaches kFp64Below, or after
// kStallChecks checks without halving the candidate kkt -- the rule of the
// single-LP engine's FP32 phase.  Measured on the 4,096-LP workload: FP32
// values take more iterations to 1e-6 (mean 2,381 against 2,229 with this
// rule; 2,882 when only a 20-check stall moves an LP) and the streamed batch
// runs at 6,842-7,049 against 7,478 LPs/s in FP64 (the FP32-value step moves
// its fewer bytes at a lower rate); switching after 3-5 checks without
// halving moves most LPs within a few checks (6,440-6,654 when the LPs on
// FP64 values ran in the tail launch).
constexpr int kStallChecks = 20;
constexpr double kFp64Below = 1e-5;

enum Status : int {
    kRunning = 0,
    kParkOptimal = 1,  // candidate with device kkt <= tol: host certifies
    kParkFarkas = 2,   // a Farkas ray could not be ruled out on the device: host checks
    kIterationLimit = 3,
    kNumerical = 4,
    kExhausted = 5,
    kHandoff = 6,      // reached the handoff iteration count: continues on the grid-wide engine
};

// Device pointers and constants of one resident LP.
struct LpDev {
    int n, m, n0, m0;
    const int *ap, *ai, *tp, *ti;  // scaled A (m rows) and A^T (n rows)
    const double *av, *tv;
    const double *c, *lb, *ub, *rl, *ru;
    const int *op, *oi, *otp, *oti;  // original A (m0 rows) and A^T (n0 rows)
    const double *ov, *otv;
    const double *c0, *lb0, *ub0, *rl0, *ru0;
    const int *cols, *lower_rows, *upper_rows;
    const double *fixed, *dc, *dr;
    double sobj, offset0, cscale;
    double *x[2], *xa[2], *dx, *y[2], *ya[2], *ax[2];
    double *epx, *epy;
    double *ox[2], *oy[2], *act, *aty, *atyabs;
    double *act1, *aty1, *atyabs1;  // the second candidate's products in a paired check
    double *bx, *by;
    const int *ablk, *tblk;  // row blocks of A and A^T (the global class's stream_rows)
    // rows longer than kChunk of A, original A, original A^T (the checks' trees)
    const int *along, *olong, *otlong;
    // A and A^T in 64-row slices for the LDS step kernel (see rows_dot_sell)
    // (entry indices in 16 bits: an LDS-class LP has m + n <= 8,000)
    const int *asb, *tsb;
    const unsigned short *asi, *tsi;
    const double *asv, *tsv;
    // the same values in FP32 at the positions of their indices (value quads
    // beside index quads), for the FP32-value step of the mixed precision
    // mode; null in FP64 mode
    const float *asv32, *tsv32;
    // (lb, ub) and (rl, ru) interleaved, so the step reads a bound pair with
    // one 16-byte load (the LDS classes; null for the global class)
    const double *lub, *rlu;
    int nab, ntb, nalong, nolong, notlong;
};

struct LpState {
    double step, weight, factor, mass, restart_kkt, best_kkt, cand_kkt;
    double first_weight;  // primal weight after the first restart
    long long iterations, since_restart, accepted, rejected, restarts, weight_updates, checks, iteration_limit;
    long long first_restart;  // iteration of the first restart (0: none yet)
    int cur, status, phase, avg;
    int consecutive_rejects, farkas_parks;
    // mixed precision: fp64 = 1 once the LP steps with FP64 values (always in
    // FP64 mode); stall_kkt / stall_checks track checks since the best
    // candidate kkt last halved
    double stall_kkt;
    long long fp32_iterations;
    int fp64, stall_checks;
};

struct BatchOpts {
    double tol;
    int check_every, adaptive, restart, mixed;
    long long handoff;  // iterations after which an LP leaves the batch (0: never)
    // k^-0.3 and k^-0.6 for k < pow_count, computed on the host with std::pow
    // (the adaptive step rule's exponents, src/cpu.cpp:156-163)
    const double *pow3, *pow6;
    long long pow_count;
};

// Beyond the tables (more than 2^22 accepted steps) the power is computed
// here; as a call of its own its constants stay out of the step kernels'
// registers.
__device__ __noinline__ double pow_beyond_table(long long k, double e) { return pow(double(k), e); }
__device__ __forceinline__ double pow_m3(const BatchOpts &o, long long k) {
    return k < o.pow_count ? o.pow3[k] : pow_beyond_table(k, -.3);
}
__device__ __forceinline__ double pow_m6(const BatchOpts &o, long long k) {
    return k < o.pow_count ? o.pow6[k] : pow_beyond_table(k, -.6);
}

// ------------------------------------------------------------ reductions --

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

__device__ __forceinline__ double maxd(double a, double b) { return a < b ? b : a; }  // std::max
__device__ __forceinline__ double mind(double a, double b) { return b < a ? b : a; }  // std::min
__device__ __forceinline__ double clampd(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }

// Reduces NS sums, NM maxima and ND double-double sums over the workgroup in
// a fixed order; results valid in every thread on return.
template <int NS, int NM, int ND>
struct Reducer {
    double s[NS > 0 ? NS : 1], mx[NM > 0 ? NM : 1];
    DD d[ND > 0 ? ND : 1];
};

// Sum (MAX: maximum) of lds[w * W + k] over waves B .. B+N-1 as a pairwise
// tree (for four waves: (w0 + w1) + (w2 + w3)).
template <int B, int N, bool MAX = false>
__device__ __forceinline__ double wave_tree(const double *lds, int W, int k) {
    if constexpr (N == 1) return lds[B * W + k];
    else {
        const double a = wave_tree<B, N / 2, MAX>(lds, W, k), b = wave_tree<B + N / 2, N / 2, MAX>(lds, W, k);
        return MAX ? maxd(a, b) : a + b;
    }
}

// v from the lane CTRL names within the lane's row of 16 lanes (a DPP row
// rotation: control 0x120 + k rotates the row right by k lanes).
template <int CTRL>
__device__ __forceinline__ double dpp_rotate(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, int(unsigned(b)), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, int(unsigned(b >> 32)), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, (unsigned long long)unsigned(hi) << 32 | unsigned(lo));
}

// One level of the wave butterfly for every value of r: each lane combines
// its value with the one `partner` brings (sums, maxima, double-double sums).
template <int NS, int NM, int ND, class Partner>
__device__ __forceinline__ void butterfly_level(Reducer<NS, NM, ND> &r, Partner partner) {
#pragma unroll
    for (int k = 0; k < NS; ++k) r.s[k] += partner(r.s[k]);
#pragma unroll
    for (int k = 0; k < NM; ++k) r.mx[k] = maxd(r.mx[k], partner(r.mx[k]));
#pragma unroll
    for (int k = 0; k < ND; ++k) r.d[k] = dd_merge(r.d[k], DD{partner(r.d[k].hi), partner(r.d[k].lo)});
}

// Workgroup reduction over the first NW waves; the lanes of any further
// waves take no part (a larger workgroup can run an NW-wave reduction and get
// exactly the NW-wave result), but every thread reaches the barriers.
// Within a wave the butterfly combines lane i with lane i ^ off for off = 32,
// 16, 8, 4, 2, 1.  The two cross-row levels are lane shuffles; the four
// within a row of 16 lanes are DPP row rotations: rotating by 8 pairs lane i
// with i ^ 8, and once the values are symmetric under ^ 8 (after that level)
// rotating by 4 brings lane i the value of lane i ^ 4, and so on for 2 and
// 1.  Every combination here is commutative bit for bit (sums; maxima of
// values that are never NaN or -0; double-double sums, whose two-sum error
// term is exact), so every lane gets exactly the butterfly's value.
template <int NW = kWaves, int NS, int NM, int ND>
__device__ void block_reduce(Reducer<NS, NM, ND> &r, double *lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (wave < NW) {
        butterfly_level(r, [](double v) { return __shfl_xor(v, 32, 64); });
        butterfly_level(r, [](double v) { return __shfl_xor(v, 16, 64); });
        butterfly_level(r, [](double v) { return dpp_rotate<0x128>(v); });  // row_ror:8
        butterfly_level(r, [](double v) { return dpp_rotate<0x124>(v); });  // row_ror:4
        butterfly_level(r, [](double v) { return dpp_rotate<0x122>(v); });  // row_ror:2
        butterfly_level(r, [](double v) { return dpp_rotate<0x121>(v); });  // row_ror:1
    }
    constexpr int W = NS + NM + 2 * ND;
    __syncthreads();
    if (lane == 0 && wave < NW) {
#pragma unroll
        for (int k = 0; k < NS; ++k) lds[wave * W + k] = r.s[k];
#pragma unroll
        for (int k = 0; k < NM; ++k) lds[wave * W + NS + k] = r.mx[k];
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            lds[wave * W + NS + NM + 2 * k] = r.d[k].hi;
            lds[wave * W + NS + NM + 2 * k + 1] = r.d[k].lo;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NS; ++k) r.s[k] = wave_tree<0, NW>(lds, W, k);
#pragma unroll
    for (int k = 0; k < NM; ++k) r.mx[k] = wave_tree<0, NW, true>(lds, W, NS + k);
#pragma unroll
    for (int k = 0; k < ND; ++k) {
        const int b = NS + NM + 2 * k;
        DD acc{lds[b], lds[b + 1]};
        for (int w = 1; w < NW; ++w) acc = dd_merge(acc, DD{lds[w * W + b], lds[w * W + b + 1]});
        r.d[k] = acc;
    }
    __syncthreads();
}

// ------------------------------------------------------------ primitives --

// CSR "stream" SpMV over consecutive row blocks of at most kChunk nonzeros:
// the workgroup computes the block's products with independent, coalesced
// loads into LDS, then each thread sums whole rows from LDS left to right
// starting at init(r) (the same order as the reference's sequential row sums,
// so the row values are bit-identical to the CPU's).  A row longer than
// kChunk forms its own block and is reduced by the workgroup in a fixed tree.
// epi(r, sum, abs_sum) consumes each row.
// LpDev's pointers are loaded from memory, so the compiler cannot tell they
// point to global memory and emits flat accesses; a flat load also counts
// against the LDS wait counter, which makes every LDS wait drain all global
// loads in flight.  Device code casts them back to global pointers.
#define GLOBAL_AS __attribute__((address_space(1)))
// A pointer that is the same in every lane, moved to scalar registers.
template <class T>
__device__ __forceinline__ T *uniform(T *p) {
    const unsigned long long v = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane(unsigned(v));
    const unsigned hi = __builtin_amdgcn_readfirstlane(unsigned(v >> 32));
    return reinterpret_cast<T *>((unsigned long long)hi << 32 | lo);
}
template <class T>
__device__ __forceinline__ const GLOBAL_AS T *gin(const T *p) {
    return (const GLOBAL_AS T *)p;
}
template <class T>
__device__ __forceinline__ GLOBAL_AS T *gout(T *p) {
    return (GLOBAL_AS T *)p;
}

constexpr int kChunk = 2048;

// Start of a thread's strided loop over [begin, end) in which only the first
// NT threads of the workgroup take part (the others start at `end`).
template <int NT>
__device__ __forceinline__ int first_index(int begin, int end) {
    return threadIdx.x < unsigned(NT) ? begin + int(threadIdx.x) : end;
}

// NT: threads of the workgroup; NTR (<= NT): threads that form the sum of a
// row longer than kChunk, so that the sum does not depend on NT.  Row blocks
// (host row_blocks): consecutive rows of at most STAGE nonzeros in all, each
// row at most kChunk long, or a single row longer than kChunk; every row up
// to kChunk long is summed in storage order by one thread whatever STAGE is.
template <bool ABS, int NT = kThreads, int NTR = NT, int STAGE = kChunk, class Init, class Epi>
__device__ void stream_rows(const int *ptr_, const int *idx_, const double *val_, const int *blk_, int nblk,
                            const double *vec_, double *stage, double *red, Init init, Epi epi) {
    const GLOBAL_AS int *ptr = gin(ptr_), *idx = gin(idx_), *blk = gin(blk_);
    const GLOBAL_AS double *val = gin(val_), *vec = gin(vec_);
    for (int b = 0; b < nblk; ++b) {
        const int r0 = blk[b], r1 = blk[b + 1];
        const int k0 = ptr[r0], k1 = ptr[r1];
        if (k1 - k0 <= kChunk || r1 - r0 > 1) {
            for (int k = k0 + threadIdx.x; k < k1; k += NT) stage[k - k0] = val[k] * vec[idx[k]];
            __syncthreads();
            for (int r = r0 + threadIdx.x; r < r1; r += NT) {
                double acc = init(r), a = 0;
                const int e = ptr[r + 1] - k0;
                for (int k = ptr[r] - k0; k < e; ++k) {
                    acc += stage[k];
                    if (ABS) a += fabs(stage[k]);
                }
                epi(r, acc, a);
            }
            __syncthreads();
        } else {
            Reducer<2, 0, 0> rr{};
            for (int k = first_index<NTR>(k0, k1); k < k1; k += NTR) {
#pragma clang fp contract(off)
                const double t = val[k] * vec[idx[k]];
                rr.s[0] += t;
                if (ABS) rr.s[1] += fabs(t);
            }
            block_reduce<NTR / 64>(rr, red);
            if (threadIdx.x == 0) epi(r0, init(r0) + rr.s[0], rr.s[1]);
            __syncthreads();
        }
    }
}

// Entries of a row that rows_direct loads at once.
constexpr int kRowGroup = 8;

// CSR row sums for the checks, one thread per row: each row of up to kChunk
// nonzeros is summed by one thread straight from memory, left to right from
// 0 over products rounded before they are added (the order of the
// reference's sequential row sums, and the values stream_rows stages), so
// the rows need no stage and no barrier between them.  The rows longer than
// kChunk, listed by the host in `longs`, are then summed one after another
// by the first NTR threads in stream_rows' tree.  NV (1 or 2) vectors share
// each entry's value and index loads; epi(r, sums, abs_sums) takes arrays of
// NV values.  Ends with a barrier.
template <bool ABS, int NV, int NT, int NTR, class Epi>
__device__ void rows_direct(const int *ptr_, const int *idx_, const double *val_, int nrows, const int *longs_,
                            int nlong, const double *vec0_, const double *vec1_, double *red, Epi epi) {
    const GLOBAL_AS int *ptr = gin(ptr_), *idx = gin(idx_);
    const GLOBAL_AS double *val = gin(val_);
    const GLOBAL_AS double *vec[2] = {gin(vec0_), gin(vec1_)};
    if constexpr (NT >= 1024) {
        // few rows per thread (the tail launch's 1,024 threads): the loads
        // of a few consecutive entries issued together
        for (int r = threadIdx.x; r < nrows; r += NT) {
            const int k0 = ptr[r], k1 = ptr[r + 1];
            if (k1 - k0 > kChunk) continue;
            double acc[NV], ab[NV];
#pragma unroll
            for (int q = 0; q < NV; ++q) acc[q] = ab[q] = 0;
#pragma unroll 4
            for (int k = k0; k < k1; ++k) {
                const double v = val[k];
                const int c = idx[k];
#pragma unroll
                for (int q = 0; q < NV; ++q) {
#pragma clang fp contract(off)
                    const double t = v * vec[q][c];
                    acc[q] += t;
                    if (ABS) ab[q] += fabs(t);
                }
            }
            epi(r, acc, ab);
        }
    } else {
        // many rows per thread (a crowded check): a thread's rows one after
        // another, the next row's offsets loaded while the current row is
        // summed; a row's entries in groups of kRowGroup whose loads are all
        // issued before their products are added (the group's positions past
        // the row end load the row's last entry and add nothing)
        int r = threadIdx.x;
        int k0 = r < nrows ? ptr[r] : 0, k1 = r < nrows ? ptr[r + 1] : 0;
        for (; r < nrows; r += NT) {
            const int rn = r + NT;
            const int k0n = rn < nrows ? ptr[rn] : 0, k1n = rn < nrows ? ptr[rn + 1] : 0;
            if (k1 - k0 <= kChunk) {
                double acc[NV], ab[NV];
#pragma unroll
                for (int q = 0; q < NV; ++q) acc[q] = ab[q] = 0;
                for (int k = k0; k < k1; k += kRowGroup) {
                    double v[kRowGroup], g[NV][kRowGroup];
                    int c[kRowGroup];
#pragma unroll
                    for (int e = 0; e < kRowGroup; ++e) {
                        const int kk = k + e < k1 ? k + e : k1 - 1;
                        v[e] = val[kk];
                        c[e] = idx[kk];
                    }
#pragma unroll
                    for (int e = 0; e < kRowGroup; ++e)
#pragma unroll
                        for (int q = 0; q < NV; ++q) g[q][e] = vec[q][c[e]];
#pragma unroll
                    for (int e = 0; e < kRowGroup; ++e)
                        if (k + e < k1) {
#pragma unroll
                            for (int q = 0; q < NV; ++q) {
#pragma clang fp contract(off)
                                const double t = v[e] * g[q][e];
                                acc[q] += t;
                                if (ABS) ab[q] += fabs(t);
                            }
                        }
                }
                epi(r, acc, ab);
            }
            k0 = k0n;
            k1 = k1n;
        }
    }
    for (int q = 0; q < nlong; ++q) {
        const int r = gin(longs_)[q], k0 = ptr[r], k1 = ptr[r + 1];
        Reducer<2 * NV, 0, 0> rr{};
        for (int k = first_index<NTR>(k0, k1); k < k1; k += NTR) {
            const double v = val[k];
            const int c = idx[k];
#pragma unroll
            for (int p = 0; p < NV; ++p) {
#pragma clang fp contract(off)
                const double t = v * vec[p][c];
                rr.s[2 * p] += t;
                if (ABS) rr.s[2 * p + 1] += fabs(t);
            }
        }
        block_reduce<NTR / 64>(rr, red);
        if (threadIdx.x == 0) {
            double v[NV], a[NV];
#pragma unroll
            for (int p = 0; p < NV; ++p) {
                v[p] = 0.0 + rr.s[2 * p];
                a[p] = rr.s[2 * p + 1];
            }
            epi(r, v, a);
        }
    }
    __syncthreads();
}

template <int NT = kThreads>
__device__ void copy_vec(double *dst, const double *src, int n) {
    GLOBAL_AS double *d = gout(dst);
    const GLOBAL_AS double *v = gin(src);
    for (int i = threadIdx.x; i < n; i += NT) d[i] = v[i];
}
template <int NT = kThreads>
__device__ void zero_vec(double *dst, int n) {
    GLOBAL_AS double *d = gout(dst);
    for (int i = threadIdx.x; i < n; i += NT) d[i] = 0;
}

// Original-space accuracy of one candidate (the non-certifying evaluate of
// src/verify.cpp:103-240 with compute_safe_bound = false) plus the device
// pre-filter of the Farkas test on its dual.
struct Acc {
    double objective, primal, dual, gap, kkt;
    int finite;
    int farkas_possible;  // the dual as a Farkas ray could not be shown non-positive
};

constexpr double kEps = 2.220446049250313e-16;

// Upper bound of min over the box of g * x for g in [lo, hi] (concave in g).
__device__ __forceinline__ double phi(double g, double l, double u) {
    if (g > 0) return isfinite(l) ? g * l : -INFINITY;
    if (g < 0) return isfinite(u) ? g * u : -INFINITY;
    return 0;
}

// Farkas pre-filter over columns for a ray whose A^T ray is in aty/atyabs.
// Adds into r: s[0] += sum of per-column upper bounds, s[1] += their |.|,
// s[2] += Lipschitz slack, mx[0] = 1 when a column bound is -inf.
template <int NTR, int NS, int NM, int ND>
__device__ void farkas_columns(const LpDev &L, Reducer<NS, NM, ND> &r) {
    for (int j = first_index<NTR>(0, L.n0); j < L.n0; j += NTR) {
        const double g = gin(L.aty)[j], ga = gin(L.atyabs)[j];
        const int len = gin(L.otp)[j + 1] - gin(L.otp)[j];
        const double e = double(len + 4) * 2 * kEps * ga;
        const double l = gin(L.lb0)[j], u = gin(L.ub0)[j];
        double ub = maxd(phi(g - e, l, u), phi(g + e, l, u));
        if (g - e <= 0 && g + e >= 0) ub = maxd(ub, 0.0);
        if (ub == -INFINITY) r.mx[0] = 1;
        else {
            r.s[0] += ub;
            r.s[1] += fabs(ub);
            const double box = maxd(isfinite(l) ? fabs(l) : 0.0, isfinite(u) ? fabs(u) : 0.0);
            r.s[2] += ga * box;
        }
    }
}

// restore_x / restore_y (src/preprocess.cpp:121-132) of both candidates of a
// check, (x0, y0) into (ox[0], oy[0]) and (x1, y1) into (ox[1], oy[1]): the
// fixed columns and zero rows first, then the scaled values scattered to
// their original positions.  Each thread loads the inputs of kRestoreBatch
// positions before it stores any (a store could alias a later load, so the
// compiler keeps them in program order).  Ends with a barrier.
constexpr int kRestoreBatch = 4;
template <int NT>
__device__ void restore_pair(const LpDev &L, const double *x0_, const double *y0_, const double *x1_,
                             const double *y1_) {
    const GLOBAL_AS double *x0 = gin(x0_), *y0 = gin(y0_), *x1 = gin(x1_), *y1 = gin(y1_);
    GLOBAL_AS double *ox0 = gout(L.ox[0]), *ox1 = gout(L.ox[1]), *oy0 = gout(L.oy[0]), *oy1 = gout(L.oy[1]);
    for (int j = threadIdx.x; j < L.n0; j += NT) {
        const double f = gin(L.fixed)[j];
        ox0[j] = f;
        ox1[j] = f;
    }
    for (int i = threadIdx.x; i < L.m0; i += NT) oy0[i] = oy1[i] = 0;
    __syncthreads();
    for (int j0 = threadIdx.x; j0 < L.n; j0 += kRestoreBatch * NT) {
        int col[kRestoreBatch];
        double a[kRestoreBatch], b[kRestoreBatch];
#pragma unroll
        for (int e = 0; e < kRestoreBatch; ++e) {
            const int j = j0 + e * NT;
            if (j < L.n) {
                const double d = gin(L.dc)[j];
                col[e] = gin(L.cols)[j];
                a[e] = x0[j] * d;
                b[e] = x1[j] * d;
            }
        }
#pragma unroll
        for (int e = 0; e < kRestoreBatch; ++e)
            if (j0 + e * NT < L.n) {
                ox0[col[e]] = a[e];
                ox1[col[e]] = b[e];
            }
    }
    for (int i0 = threadIdx.x; i0 < L.m; i0 += kRestoreBatch * NT) {
        int t0[kRestoreBatch], t1[kRestoreBatch];
        double a[kRestoreBatch], b[kRestoreBatch];
#pragma unroll
        for (int e = 0; e < kRestoreBatch; ++e) {
            const int i = i0 + e * NT;
            if (i < L.m) {
                const double r = gin(L.dr)[i], u = y0[i], v = y1[i];
                const int up = gin(L.upper_rows)[i], lo = gin(L.lower_rows)[i];
                t0[e] = u >= 0 ? up : lo;
                t1[e] = v >= 0 ? up : lo;
                a[e] = 0.0 + u * r * L.sobj;
                b[e] = 0.0 + v * r * L.sobj;
            }
        }
#pragma unroll
        for (int e = 0; e < kRestoreBatch; ++e)
            if (i0 + e * NT < L.m) {
                oy0[t0[e]] = a[e];
                oy1[t1[e]] = b[e];
            }
    }
    __syncthreads();
}

// Original-space accuracy of a restored candidate (ox, oy) whose products
// are in act = A ox, aty = A^T oy and atyabs = |A^T| |oy| (summed per entry).
// The first NTR threads form every sum.
template <int NTR>
__device__ Acc assess(const LpDev &L, const double *ox_, const double *oy_, const double *act_, const double *aty_,
                      const double *atyabs_, double *lds) {
    const GLOBAL_AS double *act = gin(act_), *aty = gin(aty_), *atyabs = gin(atyabs_);
    const GLOBAL_AS double *ox = gin(ox_), *oy = gin(oy_);
    // s : objective, Farkas column bound sum, its |.| sum, Lipschitz slack, sum |y b|
    // mx: nonfinite, Farkas column -inf, primal, primal_abs, dual, dual_abs,
    //     y != 0 at an infinite row endpoint, infinite Lagrangian box argument
    // d : support, complement, box_min, sum y b over rows
    Reducer<5, 8, 4> r{};
    // Each loop loads the next position's inputs before it works on the
    // current one, so a thread always has a position's loads in flight.
    int i = first_index<NTR>(0, L.m0);
    double na = 0, ny = 0, nlo = 0, nhi = 0;
    if (i < L.m0) na = act[i], ny = oy[i], nlo = gin(L.rl0)[i], nhi = gin(L.ru0)[i];
    for (; i < L.m0; i += NTR) {
        const double a = na, yi = ny, lo = nlo, hi = nhi;
        if (i + NTR < L.m0) {
            const int k = i + NTR;
            na = act[k], ny = oy[k], nlo = gin(L.rl0)[k], nhi = gin(L.ru0)[k];
        }
        if (!isfinite(a) || !isfinite(yi)) r.mx[0] = 1;
        const double v = maxd(maxd(0., lo - a), a - hi);
        if (isfinite(lo)) r.mx[2] = maxd(r.mx[2], maxd(0., lo - a) / (1 + fabs(lo)));
        if (isfinite(hi)) r.mx[2] = maxd(r.mx[2], maxd(0., a - hi) / (1 + fabs(hi)));
        r.mx[3] = maxd(r.mx[3], v);
        if (yi != 0) {
            const double endpoint = yi > 0 ? hi : lo;
            if (isfinite(endpoint)) {
                dd_add(r.d[0], yi * endpoint);
                dd_add(r.d[1], fabs(yi * (endpoint - a)));
                dd_add(r.d[3], yi * endpoint);
                r.s[4] += fabs(yi * endpoint);
            } else {
                r.mx[6] = 1;  // no Lagrangian bound; as a Farkas ray its value is -inf
                r.mx[4] = maxd(r.mx[4], fabs(yi));
                r.mx[5] = maxd(r.mx[5], fabs(yi));
            }
        }
    }
    int j = first_index<NTR>(0, L.n0);
    double nx = 0, nl = 0, nu = 0, nc = 0, ng = 0, nga = 0;
    int nlen = 0;
    auto load_column = [&](int k) {
        nx = ox[k], nl = gin(L.lb0)[k], nu = gin(L.ub0)[k], nc = gin(L.c0)[k];
        ng = aty[k], nga = atyabs[k], nlen = gin(L.otp)[k + 1] - gin(L.otp)[k];
    };
    if (j < L.n0) load_column(j);
    for (; j < L.n0; j += NTR) {
        const double xj = nx, l = nl, u = nu, c = nc, g = ng, ga = nga;
        const int len = nlen;
        if (j + NTR < L.n0) load_column(j + NTR);
        if (!isfinite(xj)) r.mx[0] = 1;
        const double v = maxd(maxd(0., l - xj), xj - u);
        r.mx[2] = maxd(r.mx[2], v / (1 + fabs(xj)));
        r.mx[3] = maxd(r.mx[3], v);
        r.s[0] += c * xj;
        const double linear = c + g;
        if (!isfinite(linear)) r.mx[0] = 1;
        const double d = clampd(linear, xj - u, xj - l);
        r.mx[5] = maxd(r.mx[5], fabs(d));
        r.mx[4] = maxd(r.mx[4], fabs(d) / L.cscale);
        if (linear > 0 && isfinite(l)) dd_add(r.d[1], fabs(linear * (xj - l)));
        if (linear < 0 && isfinite(u)) dd_add(r.d[1], fabs(linear * (u - xj)));
        const double arg = linear > 0 ? l : linear < 0 ? u : clampd(0., l, u);
        if (!isfinite(arg)) r.mx[7] = 1;  // no finite Lagrangian box bound
        else dd_add(r.d[2], linear * arg);
        // Farkas column bound for the candidate's dual as a ray
        const double e = double(len + 4) * 2 * kEps * ga;
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
    block_reduce<NTR / 64>(r, lds);
    Acc a{};
    const double support = dd_value(r.d[0]), complement = dd_value(r.d[1]), box_min = dd_value(r.d[2]);
    a.objective = L.offset0 + r.s[0];
    a.primal = r.mx[2];
    a.dual = r.mx[4];
    a.gap = complement / (1 + fabs(a.objective - L.offset0));
    const bool bound_ok = r.mx[6] == 0 && r.mx[7] == 0;
    if (bound_ok) {
        const double lower = (L.offset0 - support) + box_min;
        const double guard = 64 * kEps * (1 + fabs(L.offset0) + fabs(support) + fabs(box_min));
        const double lb = lower - guard;
        a.gap = maxd(a.gap, fabs(a.objective - lb) / (1 + fabs(a.objective) + fabs(lb)));
    }
    a.kkt = maxd(maxd(a.primal, a.dual), a.gap);
    a.finite = r.mx[0] == 0 && isfinite(a.objective) && isfinite(a.kkt);
    if (!a.finite) a.kkt = INFINITY;
    // Farkas: rows with y != 0 at an infinite endpoint, or a column bound of
    // -inf, make the ray's value -inf; otherwise U = sum(col bounds) - sum(y b).
    const bool neg_inf = r.mx[1] != 0 || r.mx[6] != 0;
    const double value = r.s[1] - dd_value(r.d[3]);
    const double slack = double(L.n0 + L.m0 + 16) * 4 * kEps * (r.s[2] + r.s[4] + r.s[3]);
    a.farkas_possible = !(neg_inf || value + slack <= 0);
    return a;
}


// Original-space accuracy of both candidates of a check (current and
// average): restore each into (ox, oy), then one pass over the original A and
// one over its transpose read each entry once for the two candidates'
// products, which go to (act, aty, atyabs) and (act1, aty1, atyabs1).  Each
// candidate's sums are formed as they would be alone.  NT threads copy and
// multiply; the first NTR form every sum across rows and columns.
template <int NT, int NTR>
__device__ void evaluate_pair(const LpDev &L, const double *x0, const double *y0, const double *x1, const double *y1,
                              double *lds, Acc &a0, Acc &a1) {
    restore_pair<NT>(L, x0, y0, x1, y1);
    rows_direct<false, 2, NT, NTR>(L.op, L.oi, L.ov, L.m0, L.olong, L.nolong, L.ox[0], L.ox[1], lds,
                                   [&](int i, const double *v, const double *) {
                                       gout(L.act)[i] = v[0];
                                       gout(L.act1)[i] = v[1];
                                   });
    rows_direct<true, 2, NT, NTR>(L.otp, L.oti, L.otv, L.n0, L.otlong, L.notlong, L.oy[0], L.oy[1], lds,
                                  [&](int j, const double *v, const double *a) {
                                      gout(L.aty)[j] = v[0];
                                      gout(L.atyabs)[j] = a[0];
                                      gout(L.aty1)[j] = v[1];
                                      gout(L.atyabs1)[j] = a[1];
                                  });
    a0 = assess<NTR>(L, L.ox[0], L.oy[0], L.act, L.aty, L.atyabs, lds);
    a1 = assess<NTR>(L, L.ox[1], L.oy[1], L.act1, L.aty1, L.atyabs1, lds);
}

// Accept/reject and adaptive step factor (src/cpu.cpp:151-172).  Thread 0.
// p3, p6: k^-0.3 and k^-0.6 for k = accepted + 2.
__device__ void decide_batch(LpState &S, const BatchOpts &o, double dx2, double dy2, double coupling, double tau,
                             double sigma, double f, double p3, double p6, int *accept_flag) {
    const double energy = dx2 / tau + dy2 / sigma;
    const double limit = fabs(coupling) > 0 ? energy / (2 * fabs(coupling)) : INFINITY;
    const bool finite = isfinite(energy) && isfinite(coupling);
    const bool accept = finite && (!o.adaptive || limit >= 1);
    if (o.adaptive) {
        double mult = mind((1 - p3) * limit, 1 + p6);
        if (!finite || !isfinite(mult) || mult <= 0) mult = .5;
        S.factor = clampd(f * mult, 1e-12, 1e12);
    }
    if (accept) {
        S.cur ^= 1;
        S.accepted += 1;
        S.mass += f;
        S.consecutive_rejects = 0;
    } else {
        S.rejected += 1;
        S.consecutive_rejects += 1;
    }
    *accept_flag = accept ? 1 : (o.adaptive && S.consecutive_rejects < kMaxRejects ? 0 : -1);
}

__device__ void attempt(const LpDev &L, LpState &S, const BatchOpts &o, double tau0, double sigma0, double *lds,
                        double *stage, int *accept_flag) {
    const int cur = S.cur;
    const double f = S.factor, tau = tau0 * f, sigma = sigma0 * f, w = f / (S.mass + f);
    const double *x = L.x[cur], *xa = L.xa[cur], *y = L.y[cur], *ya = L.ya[cur], *ax = L.ax[cur];
    double *xn = L.x[cur ^ 1], *xan = L.xa[cur ^ 1], *yn = L.y[cur ^ 1], *yan = L.ya[cur ^ 1], *axn = L.ax[cur ^ 1];
    Reducer<3, 0, 0> r{};
    // g = c + A^T y and the projected primal step (src/cpu.cpp:107-132)
    stream_rows<false>(L.tp, L.ti, L.tv, L.tblk, L.ntb, y, stage, lds, [&](int j) { return gin(L.c)[j]; },
                       [&](int j, double g, double) {
                           const double xj = x[j];
                           const double t = clampd(xj - tau * g, gin(L.lb)[j], gin(L.ub)[j]);
                           const double d = t - xj;
                           r.s[0] += d * d;
                           L.dx[j] = d;
                           xn[j] = t;
                           const double v = xa[j];
                           xan[j] = v + (t - v) * w;
                       });
    // delta = A dx and the extrapolated dual step (src/cpu.cpp:136-150)
    stream_rows<false>(L.ap, L.ai, L.av, L.ablk, L.nab, L.dx, stage, lds, [](int) { return 0.0; },
                       [&](int i, double delta, double) {
                           const double a = ax[i], bar = a + 2 * delta, yi = y[i], z = yi + sigma * bar;
                           const double t = maxd(0., z - sigma * gin(L.ru)[i]) + mind(0., z - sigma * gin(L.rl)[i]);
                           const double dy = t - yi;
                           r.s[1] += (!isfinite(bar) || !isfinite(t)) ? INFINITY : dy * dy;
                           r.s[2] += dy * delta;
                           yn[i] = t;
                           const double v = ya[i];
                           yan[i] = v + (t - v) * w;
                           axn[i] = a + delta;
                       });
    block_reduce(r, lds);
    if (threadIdx.x == 0)
        decide_batch(S, o, r.s[0], r.s[1], r.s[2], tau, sigma, f, pow_m3(o, S.accepted + 2), pow_m6(o, S.accepted + 2),
                     accept_flag);
    __syncthreads();
}

// Phase of a resident LP between the two kernels.
enum Phase : int { kAdvance = 0, kResume = 1, kCheck = 2 };

// Hot loop: `check_every` accepted steps for every running LP whose next
// action is to advance (src/cpu.cpp:103-220 per step).
__global__ __launch_bounds__(kThreads) void batch_advance_kernel(const LpDev *lps, LpState *states, const int *slots,
                                                                 BatchOpts o) {
    const int s = slots[blockIdx.x];
    __shared__ LpState S;
    __shared__ double lds[kWaves * 4];
    __shared__ double stage[kChunk];
    __shared__ int flag;
    if (threadIdx.x == 0) S = states[s];
    __syncthreads();
    if (S.status != kRunning || S.phase != kAdvance) return;
    const LpDev &L = lps[s];
    const long long remaining = S.iteration_limit - S.iterations;
    if (remaining <= 0) {
        if (threadIdx.x == 0) states[s].status = kIterationLimit;
        return;
    }
    const long long count = remaining < o.check_every ? remaining : o.check_every;
    const double tau0 = S.step / S.weight, sigma0 = S.step * S.weight;
    for (long long done = 0; done < count;) {
        attempt(L, S, o, tau0, sigma0, lds, stage, &flag);
        if (flag == 1) ++done;
        else if (flag < 0) {
            if (threadIdx.x == 0) S.status = kExhausted;
            break;
        }
    }
    if (threadIdx.x == 0) {
        if (S.status == kRunning) {
            S.iterations += count;
            S.since_restart += count;
            S.checks += 1;
            S.phase = kCheck;
        }
        states[s] = S;
    }
}

// Sequential dot product of one CSR row against an LDS-resident vector, in
// storage order starting from `init` (the reference's row-sum order), with
// four entries' loads issued together.
__device__ __forceinline__ double row_dot(const double *val, const int *idx, int k, int end, const double *vec,
                                          double acc) {
    for (; k + 4 <= end; k += 4) {
        const double v0 = val[k], v1 = val[k + 1], v2 = val[k + 2], v3 = val[k + 3];
        const int i0 = idx[k], i1 = idx[k + 1], i2 = idx[k + 2], i3 = idx[k + 3];
        acc += v0 * vec[i0];
        acc += v1 * vec[i1];
        acc += v2 * vec[i2];
        acc += v3 * vec[i3];
    }
    for (; k < end; ++k) acc += val[k] * vec[idx[k]];
    return acc;
}

// Dot products of NR CSR rows (entries [k[q], e[q]); absent rows have
// k == e) against an LDS-resident vector.  The rows advance in lock step,
// kStep entries per row per step, so up to NR * kStep independent loads are in
// flight per thread (a row of at most kStep entries costs one round trip);
// each row is still summed in storage order starting from acc[q].
constexpr int kStep = 8;

template <int NR>
__device__ __forceinline__ void rows_dot(const int *k, const int *e, const int *idx, const double *val,
                                         const double *vec, double *acc) {
    int longest = 0;
#pragma unroll
    for (int q = 0; q < NR; ++q) longest = max(longest, e[q] - k[q]);
    for (int t = 0; t < longest; t += kStep) {
        double v[NR][kStep];
        int ix[NR][kStep];
#pragma unroll
        for (int q = 0; q < NR; ++q)
#pragma unroll
            for (int u = 0; u < kStep; ++u) {
                const int kk = k[q] + t + u;
                const bool on = kk < e[q];
                v[q][u] = on ? val[kk] : 0.0;
                ix[q][u] = on ? idx[kk] : 0;
            }
#pragma unroll
        for (int q = 0; q < NR; ++q)
#pragma unroll
            for (int u = 0; u < kStep; ++u)
                if (k[q] + t + u < e[q]) acc[q] += v[q][u] * vec[ix[q][u]];
    }
}

// Dot products like rows_dot, read from the sliced layout: the rows of A (or
// A^T) are cut into slices of 64 consecutive rows, one per wave, padded to an
// even width w at least the slice's longest row.  Values are stored in
// pairs, entry-major (entries 2t and 2t+1 of the 64 rows are 128 consecutive
// values, lane l's two side by side), indices in quads (entries 4t..4t+3 of
// the 64 rows are 256 consecutive 16-bit indices, lane l's four side by
// side; when w is not a multiple of 4 the last two entries form a pair
// region of 128).  A wave's load of a value pair is one contiguous 1 KiB
// request and of four indices one 512-byte request: the step kernel is bound
// by the CU's rate of memory instructions, not by bytes, and the grouping
// takes the matrix from two instructions per entry to three per four.  Each
// row is still summed in storage order, so the sums equal rows_dot's.
typedef double double2v __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));
typedef float float4v __attribute__((ext_vector_type(4)));
typedef unsigned short ushort2v __attribute__((ext_vector_type(2)));
typedef unsigned short ushort4v __attribute__((ext_vector_type(4)));
// VT = double: FP64 values in pairs.  VT = float (mixed precision): FP32
// values stored like the indices, four entries of a lane per 16-byte load
// (a last pair region for widths that are not a multiple of 4); products and
// sums stay FP64.
template <int NR, class VT = double>
__device__ __forceinline__ void rows_dot_sell(const int *rlen, const int *sb,
                                              const GLOBAL_AS unsigned short *si, const GLOBAL_AS VT *sv, int slices, int b0,
                                              int nt, const double *vec, double *acc) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // Every lane of a slice owns `w` stored entries (padding is value 0,
    // index 0), so loads below a slice's width need no per-lane test: only
    // the sum skips a row's padding.  Widths are the same for the whole wave.
    int base[NR], len[NR], w[NR], width = 0, common = 1 << 30;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        const int s = (b0 + q * nt) / 64 + wave;
        base[q] = 0;
        w[q] = 0;
        len[q] = rlen[q];
        if (s < slices) {
            const int b = sb[s];
            base[q] = b;
            w[q] = (sb[s + 1] - b) / 64;
        }
        base[q] = __builtin_amdgcn_readfirstlane(base[q]);
        w[q] = __builtin_amdgcn_readfirstlane(w[q]);
        width = max(width, w[q]);
        common = min(common, w[q]);
    }
    static_assert(kStep % 4 == 0, "chunks hold whole index quads");
    // one chunk of kStep entries per row; `full` chunks lie within every
    // row's slice and need no width test
    auto chunk = [&](int t, bool full) {
        double v[NR][kStep];
        int ix[NR][kStep];
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            if constexpr (sizeof(VT) == 8) {
                // value pair t/2 + h of this lane at base + (t/2 + h) * 128 + 2 lane
                const GLOBAL_AS double2v *pv = (const GLOBAL_AS double2v *)(sv + base[q] + t * 64 + 2 * lane);
#pragma unroll
                for (int h = 0; h < kStep / 2; ++h) {
                    if (full || t + 2 * h < w[q]) {
                        const double2v vv = pv[h * 64];
                        v[q][2 * h] = vv.x;
                        v[q][2 * h + 1] = vv.y;
                    } else {
                        v[q][2 * h] = v[q][2 * h + 1] = 0.0;
                    }
                }
            } else {
                // value quad t/4 + h at base + (t/4 + h) * 256 + 4 lane, as the indices
                const GLOBAL_AS float *pf = sv + base[q] + t * 64;
#pragma unroll
                for (int h = 0; h < kStep / 4; ++h) {
                    const int u0 = t + 4 * h;
                    float4v f = {0.f, 0.f, 0.f, 0.f};
                    if (full || u0 + 4 <= w[q]) {
                        f = *(const GLOBAL_AS float4v *)(pf + h * 256 + 4 * lane);
                    } else if (u0 + 2 <= w[q]) {
                        const float2v f2 = *(const GLOBAL_AS float2v *)(pf + h * 256 + 2 * lane);
                        f.x = f2.x;
                        f.y = f2.y;
                    }
                    v[q][4 * h] = double(f.x);
                    v[q][4 * h + 1] = double(f.y);
                    v[q][4 * h + 2] = double(f.z);
                    v[q][4 * h + 3] = double(f.w);
                }
            }
            // index quad t/4 + h at base + (t/4 + h) * 256 + 4 lane (a last
            // pair region at + 2 lane)
            const GLOBAL_AS unsigned short *pi = si + base[q] + t * 64;
#pragma unroll
            for (int h = 0; h < kStep / 4; ++h) {
                const int u0 = t + 4 * h;
                ushort4v ii = {0, 0, 0, 0};
                if (full || u0 + 4 <= w[q]) {
                    ii = *(const GLOBAL_AS ushort4v *)(pi + h * 256 + 4 * lane);
                } else if (u0 + 2 <= w[q]) {
                    const ushort2v p2 = *(const GLOBAL_AS ushort2v *)(pi + h * 256 + 2 * lane);
                    ii.x = p2.x;
                    ii.y = p2.y;
                }
                ix[q][4 * h] = ii.x;
                ix[q][4 * h + 1] = ii.y;
                ix[q][4 * h + 2] = ii.z;
                ix[q][4 * h + 3] = ii.w;
            }
        }
#pragma unroll
        for (int q = 0; q < NR; ++q)
#pragma unroll
            for (int u = 0; u < kStep; ++u)
                if (full || t + u < w[q]) {
                    const double p = v[q][u] * vec[ix[q][u]];
                    acc[q] = t + u < len[q] ? acc[q] + p : acc[q];
                }
    };
    int t = 0;
    for (; t + kStep <= common; t += kStep) chunk(t, true);
    for (; t < width; t += kStep) chunk(t, false);
}

// Sums over rows are taken per tile of 64 consecutive rows (one wave, fixed
// butterfly) and the tiles are then added in tile order, so an attempt's energy
// and coupling do not depend on the workgroup size or rows per thread: an LP
// can change launch shape (the tail mode below) and still follow exactly the
// trajectory it follows alone.
constexpr int kMaxTiles = 128;  // 64-row tiles: m, n <= 8000 in the LDS classes

// Sum of v over the 64 lanes of a wave, valid in every lane: the butterfly
// of block_reduce (lane shuffles across rows, DPP row rotations within
// them), so the sum is the butterfly's bit for bit.
__device__ __forceinline__ double wave_sum(double v) {
    v += __shfl_xor(v, 32, 64);
    v += __shfl_xor(v, 16, 64);
    v += dpp_rotate<0x128>(v);  // row_ror:8
    v += dpp_rotate<0x124>(v);  // row_ror:4
    v += dpp_rotate<0x122>(v);  // row_ror:2
    v += dpp_rotate<0x121>(v);  // row_ror:1
    return v;
}

// Sum of tiles[0..count) in tile order, in a fixed tree over one wave; the
// result is valid in every lane of wave 0.
__device__ __forceinline__ double tile_total(const double *tiles, int count) {
    const int lane = threadIdx.x & 63;
    double acc = 0;
    for (int t = lane; t < count; t += 64) acc += tiles[t];
    return wave_sum(acc);
}

// The sliced values a step of value type VT reads.
template <class VT>
__device__ __forceinline__ const GLOBAL_AS VT *sliced_values(const LpDev &L, bool transpose) {
    if constexpr (sizeof(VT) == 8) return gin(transpose ? L.tsv : L.asv);
    else return gin(transpose ? L.tsv32 : L.asv32);
}

// An LP's dynamic LDS in the step kernels: y (m doubles) and dx (n doubles),
// the slice bases of A^T and A (ceil(n/64) + 1 and ceil(m/64) + 1 ints), and
// the row lengths of A^T and A (n and m bytes: an LDS-class row holds at most
// kShortRow entries).
__device__ __forceinline__ void lds_layout(double *dyn, const LpDev &L, double *&ylds, double *&dxlds, int *&tsblds,
                                           int *&asblds, unsigned char *&tlen, unsigned char *&alen) {
    const int ns = (L.n + 63) / 64, ms = (L.m + 63) / 64;
    ylds = dyn;
    dxlds = dyn + L.m;
    tsblds = reinterpret_cast<int *>(dyn + L.m + L.n);
    asblds = tsblds + ns + 1;
    tlen = reinterpret_cast<unsigned char *>(asblds + ms + 1);
    alen = tlen + L.n;
}
template <int NT>
__device__ void stage_layout(const LpDev &L, int *tsblds, int *asblds, unsigned char *tlen, unsigned char *alen) {
    const int ns = (L.n + 63) / 64, ms = (L.m + 63) / 64;
    const GLOBAL_AS int *tp = gin(L.tp), *ap = gin(L.ap), *tsb = gin(L.tsb), *asb = gin(L.asb);
    for (int j = threadIdx.x; j < L.n; j += NT) tlen[j] = (unsigned char)(tp[j + 1] - tp[j]);
    for (int i = threadIdx.x; i < L.m; i += NT) alen[i] = (unsigned char)(ap[i + 1] - ap[i]);
    for (int j = threadIdx.x; j <= ns; j += NT) tsblds[j] = tsb[j];
    for (int i = threadIdx.x; i <= ms; i += NT) asblds[i] = asb[i];
}

// One PDHG attempt for an LP whose dual iterate and primal step fit in LDS:
// y lives in LDS (staged from memory only when `ystaged` is false: at the
// start, after a rejected attempt, after a check; an accepted attempt leaves
// its new y there), dx never leaves LDS, the row lengths and slice bases of
// A and A^T sit in LDS for the whole launch (so a row's entries are one
// memory round trip away, not two), and each thread carries NR rows at a
// time (rows are short in this class).
template <int NT, int NR, class VT = double>
__device__ void attempt_lds(const LpDev &L, LpState &S, const BatchOpts &o, double tau0, double sigma0,
                            double *tiles, double *ylds, double *dxlds, const unsigned char *tlen,
                            const unsigned char *alen, const int *tsblds, const int *asblds, double *powlds,
                            int *accept_flag,
                            bool &ystaged) {
    const int cur = S.cur;
    const double f = S.factor, tau = tau0 * f, sigma = sigma0 * f, w = f / (S.mass + f);
    // (selects rather than indexing by cur, so that a descriptor held in
    // registers stays in registers)
    const bool c1 = cur != 0;
    const GLOBAL_AS double *x = gin(c1 ? L.x[1] : L.x[0]), *xa = gin(c1 ? L.xa[1] : L.xa[0]),
                           *y = gin(c1 ? L.y[1] : L.y[0]), *ya = gin(c1 ? L.ya[1] : L.ya[0]),
                           *ax = gin(c1 ? L.ax[1] : L.ax[0]);
    GLOBAL_AS double *xn = gout(c1 ? L.x[0] : L.x[1]), *xan = gout(c1 ? L.xa[0] : L.xa[1]),
                     *yn = gout(c1 ? L.y[0] : L.y[1]), *yan = gout(c1 ? L.ya[0] : L.ya[1]),
                     *axn = gout(c1 ? L.ax[0] : L.ax[1]);
    const GLOBAL_AS double *c = gin(L.c);
    const GLOBAL_AS double2v *lub = (const GLOBAL_AS double2v *)L.lub, *rlu = (const GLOBAL_AS double2v *)L.rlu;
    // the step rule's powers load with y, off the decision's serial path
    // (the barriers after the two passes order them before the decision)
    if (threadIdx.x == 64 % NT) powlds[0] = pow_m3(o, S.accepted + 2);
    if (threadIdx.x == 65 % NT) powlds[1] = pow_m6(o, S.accepted + 2);
    if (!ystaged) {
        for (int i = threadIdx.x; i < L.m; i += NT) ylds[i] = y[i];
        __syncthreads();
    }
    double *tile_dx = tiles, *tile_dy = tiles + kMaxTiles, *tile_cr = tiles + 2 * kMaxTiles;
    const int wave = threadIdx.x >> 6;
    // g = c + A^T y and the projected primal step (src/cpu.cpp:107-132)
    // Passes are uniform over the workgroup (the tile sums need whole waves).
    for (int b0 = 0; b0 < L.n; b0 += NR * NT) {
        const int base = b0 + int(threadIdx.x);
        int r[NR], len[NR];
        bool ok[NR];
        double g[NR], xj_[NR], xa_[NR], lo_[NR], hi_[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            r[q] = base + q * NT;
            ok[q] = r[q] < L.n;
            const int j = ok[q] ? r[q] : 0;
            len[q] = ok[q] ? tlen[j] : 0;
            g[q] = ok[q] ? c[j] : 0.0;
            xj_[q] = x[j];
            xa_[q] = xa[j];
            const double2v bj = lub[j];
            lo_[q] = bj.x;
            hi_[q] = bj.y;
        }
        rows_dot_sell<NR, VT>(len, tsblds, gin(L.tsi), sliced_values<VT>(L, true), (L.n + 63) / 64, b0, NT, ylds, g);
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            double d2 = 0;
            if (ok[q]) {
                const int j = r[q];
                const double xj = xj_[q], v = xa_[q];
                const double t = clampd(xj - tau * g[q], lo_[q], hi_[q]);
                const double d = t - xj;
                d2 = d * d;
                dxlds[j] = d;
                xn[j] = t;
                xan[j] = v + (t - v) * w;
            }
            const int tile = (b0 + q * NT) / 64 + wave;
            d2 = wave_sum(d2);
            if ((threadIdx.x & 63) == 0 && tile * 64 < L.n) tile_dx[tile] = d2;
        }
    }
    __syncthreads();
    // delta = A dx and the extrapolated dual step (src/cpu.cpp:136-150)
    for (int b0 = 0; b0 < L.m; b0 += NR * NT) {
        const int base = b0 + int(threadIdx.x);
        int r[NR], len[NR];
        bool ok[NR];
        double delta[NR], a_[NR], ya_[NR], lo_[NR], hi_[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            r[q] = base + q * NT;
            ok[q] = r[q] < L.m;
            const int i = ok[q] ? r[q] : 0;
            len[q] = ok[q] ? alen[i] : 0;
            delta[q] = 0.0;
            a_[q] = ax[i];
            ya_[q] = ya[i];
            const double2v bi = rlu[i];
            lo_[q] = bi.x;
            hi_[q] = bi.y;
        }
        rows_dot_sell<NR, VT>(len, asblds, gin(L.asi), sliced_values<VT>(L, false), (L.m + 63) / 64, b0, NT, dxlds, delta);
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            double sq = 0, cr = 0;
            if (ok[q]) {
                const int i = r[q];
                const double a = a_[q], yi = ylds[i], v = ya_[q];
                const double bar = a + 2 * delta[q], z = yi + sigma * bar;
                const double t = maxd(0., z - sigma * hi_[q]) + mind(0., z - sigma * lo_[q]);
                const double dy = t - yi;
                sq = (!isfinite(bar) || !isfinite(t)) ? INFINITY : dy * dy;
                cr = dy * delta[q];
                yn[i] = t;
                ylds[i] = t;  // only this thread reads row i in this pass
                yan[i] = v + (t - v) * w;
                axn[i] = a + delta[q];
            }
            const int tile = (b0 + q * NT) / 64 + wave;
            sq = wave_sum(sq);
            cr = wave_sum(cr);
            if ((threadIdx.x & 63) == 0 && tile * 64 < L.m) {
                tile_dy[tile] = sq;
                tile_cr[tile] = cr;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const double dx2 = tile_total(tile_dx, (L.n + 63) / 64);
        const double dy2 = tile_total(tile_dy, (L.m + 63) / 64);
        const double cross = tile_total(tile_cr, (L.m + 63) / 64);
        if (threadIdx.x == 0) decide_batch(S, o, dx2, dy2, cross, tau, sigma, f, powlds[0], powlds[1], accept_flag);
    }
    __syncthreads();
    // accepted: LDS holds the new y[cur]; otherwise y[cur] is still the old one
    ystaged = *accept_flag == 1;
}

// Hot loop for the LDS-resident class; dynamic LDS holds y (m) and dx (n).
// VT: the value type of the step (float: the FP32-value phase of the mixed
// precision mode; a class's LPs that moved to FP64 values step in the same
// period's VT = double launch).
template <int NT, int NR, class VT = double>
__global__ __launch_bounds__(NT) void batch_advance_lds_kernel(const LpDev *lps, LpState *states, const int *slots,
                                                               BatchOpts o) {
    extern __shared__ double dyn[];
    const int s = slots[blockIdx.x];
    __shared__ LpState S;
    __shared__ double tiles[3 * kMaxTiles];
    __shared__ double powlds[2];
    __shared__ int flag;
    if (threadIdx.x == 0) S = states[s];
    __syncthreads();
    // an LP whose check moved it to FP64 values in this round skips the rest
    // of the round's FP32-value launches (the next round lists it for the
    // FP64-value launches), so its periods do not depend on the launch it is in
    if (S.status != kRunning || S.phase != kAdvance || (sizeof(VT) == 4) == (S.fp64 != 0)) return;
    const LpDev &L = lps[s];
    const long long remaining = S.iteration_limit - S.iterations;
    if (remaining <= 0) {
        if (threadIdx.x == 0) states[s].status = kIterationLimit;
        return;
    }
    const long long count = remaining < o.check_every ? remaining : o.check_every;
    const double tau0 = S.step / S.weight, sigma0 = S.step * S.weight;
    double *ylds, *dxlds;
    int *tsblds, *asblds;
    unsigned char *tlen, *alen;
    lds_layout(dyn, L, ylds, dxlds, tsblds, asblds, tlen, alen);
    stage_layout<NT>(L, tsblds, asblds, tlen, alen);
    __syncthreads();
    bool ystaged = false;
    for (long long done = 0; done < count;) {
        attempt_lds<NT, NR, VT>(L, S, o, tau0, sigma0, tiles, ylds, dxlds, tlen, alen, tsblds, asblds, powlds, &flag,
                                ystaged);
        if (flag == 1) ++done;
        else if (flag < 0) {
            if (threadIdx.x == 0) S.status = kExhausted;
            break;
        }
    }
    if (threadIdx.x == 0) {
        if (S.status == kRunning) {
            S.iterations += count;
            S.since_restart += count;
            S.checks += 1;
            S.phase = kCheck;
            if (sizeof(VT) == 4) S.fp32_iterations += count;
        }
        states[s] = S;
    }
}

// Many small device-to-host copies at once: piece b (source, offset, count)
// is copied into one staging buffer, which then travels in a single copy.
struct CopyPiece {
    const double *src;
    long long offset, count;
};

__global__ void gather_pieces_kernel(const CopyPiece *pieces, double *out) {
    const CopyPiece p = pieces[blockIdx.x];
    for (long long i = threadIdx.x; i < p.count; i += blockDim.x) out[p.offset + i] = p.src[i];
}

// Check period end (src/solver.cpp:550-725): candidate evaluation, best
// tracking, parking for host certification / Farkas checks, restart rule.
// For an LP whose phase is kCheck or kResume; S is the workgroup's shared
// copy of its state.  NT threads copy and multiply, the first kThreads form
// every sum across rows and columns, so the check decides the same in any
// workgroup size.  lds holds kWaves * 24 doubles.

// ax = A x of the current iterate in the row order of the scaled A.
template <int NT, int NTR>
__device__ void rebase_activity(const LpDev &L, int cur, double *lds) {
    rows_direct<false, 1, NT, NTR>(L.ap, L.ai, L.av, L.m, L.along, L.nalong, L.x[cur], nullptr, lds,
                                   [&](int r, const double *v, const double *) { gout(L.ax[cur])[r] = v[0]; });
}

template <int NT>
__device__ void check_period(const LpDev &L, LpState &S, const BatchOpts &o, double *lds) {
    constexpr int NTR = kThreads;
    const int cur = S.cur;
    double cand_kkt;
    int avg;
    if (S.phase == kCheck) {
        // candidates(): rebase the activity recurrence (src/cpu.cpp:229-230)
        rebase_activity<NT, NTR>(L, cur, lds);
        Acc cur_acc, avg_acc;
        evaluate_pair<NT, NTR>(L, L.x[cur], L.y[cur], L.xa[cur], L.ya[cur], lds, cur_acc, avg_acc);
        avg = avg_acc.kkt < cur_acc.kkt ? 1 : 0;
        const Acc ca = avg ? avg_acc : cur_acc;
        const bool better = ca.finite && ca.kkt < S.best_kkt;
        if (better) {
            copy_vec<NT>(L.bx, L.ox[avg], L.n0);
            copy_vec<NT>(L.by, L.oy[avg], L.m0);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            if (better) S.best_kkt = ca.kkt;
            S.avg = avg;
            S.cand_kkt = ca.kkt;
            S.phase = kAdvance;
            if (!ca.finite) S.status = kNumerical;
            else if (ca.kkt <= o.tol) S.status = kParkOptimal;
            // mixed precision: FP64 values from the check on whose
            // candidate reaches kFp64Below or after which kStallChecks checks
            // passed without halving the candidate kkt (the FP32 values'
            // rounding bounds how close the iterates get)
            if (o.mixed && !S.fp64) {
                if (ca.kkt < .5 * S.stall_kkt) {
                    S.stall_kkt = ca.kkt;
                    S.stall_checks = 0;
                } else if (++S.stall_checks >= kStallChecks) {
                    S.fp64 = 1;
                }
                if (ca.kkt <= kFp64Below) S.fp64 = 1;
            }
        }
        __syncthreads();
        if (S.status != kRunning) return;
        if (S.iterations >= 1000 && ca.primal > o.tol) {
            bool possible = cur_acc.farkas_possible || avg_acc.farkas_possible;
            if (!possible) {
                // displacement ray restore_y(cy - epochy) (src/solver.cpp:603-610)
                const double *cy = avg ? L.ya[cur] : L.y[cur];
                zero_vec<NT>(L.oy[0], L.m0);
                __syncthreads();
                for (int i = threadIdx.x; i < L.m; i += NT) {
                    const double d = cy[i] - gin(L.epy)[i];
                    const int t = d >= 0 ? gin(L.upper_rows)[i] : gin(L.lower_rows)[i];
                    gout(L.oy[0])[t] = 0.0 + d * gin(L.dr)[i] * L.sobj;
                }
                __syncthreads();
                rows_direct<true, 1, NT, NTR>(L.otp, L.oti, L.otv, L.n0, L.otlong, L.notlong, L.oy[0], nullptr,
                                              lds, [&](int j, const double *v, const double *a) {
                                                  gout(L.aty)[j] = v[0];
                                                  gout(L.atyabs)[j] = a[0];
                                              });
                Reducer<4, 2, 1> r{};
                for (int i = first_index<NTR>(0, L.m0); i < L.m0; i += NTR) {
                    const double yi = L.oy[0][i];
                    if (!isfinite(yi)) r.mx[1] = 1;
                    if (yi != 0) {
                        const double e = yi > 0 ? gin(L.ru0)[i] : gin(L.rl0)[i];
                        if (!isfinite(e)) r.mx[0] = 1;
                        else {
                            dd_add(r.d[0], yi * e);
                            r.s[3] += fabs(yi * e);
                        }
                    }
                }
                farkas_columns<NTR>(L, r);
                block_reduce<NTR / 64>(r, lds);
                const double value = r.s[0] - dd_value(r.d[0]);
                const double slack = double(L.n0 + L.m0 + 16) * 4 * kEps * (r.s[1] + r.s[2] + r.s[3]);
                // all-zero and nonfinite rays are skipped by the reference
                possible = !(r.mx[0] != 0 || r.mx[1] != 0 || value + slack <= 0);
            }
            if (possible) {
                __syncthreads();
                if (threadIdx.x == 0) {
                    S.status = kParkFarkas;
                    S.farkas_parks += 1;
                }
                __syncthreads();
                return;
            }
        }
        cand_kkt = ca.kkt;
    } else {
        // resume after a host check that did not finish the LP
        cand_kkt = S.cand_kkt;
        avg = S.avg;
        __syncthreads();
        if (threadIdx.x == 0) S.phase = kAdvance;
    }
    // restart rule and primal weight (src/solver.cpp:692-725)
    const bool due = S.since_restart >= 2LL * o.check_every &&
                     (cand_kkt < .5 * S.restart_kkt ||
                      S.since_restart >= (S.iterations / 2 > 2000 ? S.iterations / 2 : 2000));
    if (o.restart && due) {
        const double *cx = avg ? L.xa[cur] : L.x[cur];
        const double *cy = avg ? L.ya[cur] : L.y[cur];
        Reducer<0, 0, 2> r{};
        for (int j = first_index<NTR>(0, L.n); j < L.n; j += NTR) {
            const double d = cx[j] - gin(L.epx)[j];
            dd_add(r.d[0], d * d);
        }
        for (int i = first_index<NTR>(0, L.m); i < L.m; i += NTR) {
            const double d = cy[i] - gin(L.epy)[i];
            dd_add(r.d[1], d * d);
        }
        block_reduce<NTR / 64>(r, lds);
        if (threadIdx.x == 0 && o.adaptive) {
            const double dx = dd_value(r.d[0]), dy = dd_value(r.d[1]);
            if (dx > 1e-24 && dy > 1e-24) {
                S.weight = sqrt(S.weight * clampd(sqrt(dy / dx), 1e-4, 1e4));
                S.weight_updates += 1;
            }
        }
        if (avg) {
            copy_vec<NT>(L.x[cur], L.xa[cur], L.n);
            copy_vec<NT>(L.y[cur], L.ya[cur], L.m);
            __syncthreads();
        }
        rebase_activity<NT, NTR>(L, cur, lds);
        copy_vec<NT>(L.epx, L.x[cur], L.n);
        copy_vec<NT>(L.epy, L.y[cur], L.m);
        zero_vec<NT>(L.xa[cur], L.n);
        zero_vec<NT>(L.ya[cur], L.m);
        __syncthreads();
        if (threadIdx.x == 0) {
            S.mass = 0;
            S.restart_kkt = cand_kkt;
            S.since_restart = 0;
            S.restarts += 1;
            if (S.restarts == 1) {
                S.first_restart = S.iterations;
                S.first_weight = S.weight;
            }
        }
    }
    __syncthreads();
    // A long-running LP leaves the batch at a check, so the point where it
    // moves depends only on its own iteration count.
    if (threadIdx.x == 0 && o.handoff > 0 && S.status == kRunning && S.iterations >= o.handoff) S.status = kHandoff;
    __syncthreads();
}

// check_period as a call of its own for the fused tail kernel, so that the
// check's register needs do not shape the register allocation of the
// attempts.  The pointers carry their address spaces (LDS, global) into the
// call, so its LDS and global accesses stay direct.
#define LDS_AS __attribute__((address_space(3)))
template <int NT>
__device__ __noinline__ void check_period_call(const GLOBAL_AS LpDev *L, LDS_AS LpState *S, BatchOpts o,
                                               LDS_AS double *lds) {
    check_period<NT>(*(const LpDev *)L, *(LpState *)S, o, (double *)lds);
}

// Held to 4 waves per SIMD (128 VGPRs, some spills): at its natural 248 VGPRs a
// check workgroup leaves no room on its CU for a 1,024-thread step workgroup
// of another class until it ends (streamed bench +1.5 %, results unchanged).
__global__ __launch_bounds__(kThreads, 4) void batch_check_kernel(const LpDev *lps, LpState *states, const int *slots,
                                                               BatchOpts o) {
    const int s = slots[blockIdx.x];
    __shared__ LpState S;
    __shared__ double lds[kWaves * 24];
    if (threadIdx.x == 0) S = states[s];
    __syncthreads();
    if (S.status != kRunning || S.phase == kAdvance) return;
    check_period<kThreads>(lps[s], S, o, lds);
    if (threadIdx.x == 0) states[s] = S;
}

// The tail group in one launch, as a persistent grid over a device-side work
// queue: the launch has at most one workgroup per CU, and each workgroup
// takes the next LP of the group from the queue (an atomic counter over the
// group's slot list), drives it through whole check periods -- the attempts
// of batch_advance_lds_kernel followed by check_period -- until the LP stops
// running (parked for the host, finished, handed off), `max_periods` periods
// are done, or at least `min_periods` are done and `budget` ticks of the
// constant-rate clock have passed since the workgroup started; then it takes
// the next LP, until the queue is empty.  An LP's periods do not wait for the
// other LPs of the launch, an LP that stops early frees its workgroup for the
// next one, and a group larger than the grid shares one budget window instead
// of spending it once per wave of workgroups.  Each LP's trajectory is the
// one the two-kernel cycle gives: the attempts are the same code, and
// check_period forms its sums on the first kThreads threads exactly as
// batch_check_kernel does.  `queue` is zero at launch.
template <int NT>
__global__ __launch_bounds__(NT) void batch_tail_kernel(const LpDev *lps, LpState *states, const int *slots,
                                                        int count, unsigned *queue, BatchOpts o, int min_periods,
                                                        int max_periods, long long budget) {
    extern __shared__ double dyn[];
    const long long start = wall_clock64();
    __shared__ LpState S;
    __shared__ double tiles[3 * kMaxTiles];
    __shared__ double powlds[2];
    __shared__ double lds[kWaves * 24];
    __shared__ int flag, stop, taken;
    for (;;) {
        __syncthreads();  // the previous LP's readers of S and `taken` are done
        if (threadIdx.x == 0)
            taken = int(__hip_atomic_fetch_add(queue, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        __syncthreads();
        const int q = taken;
        if (q >= count) break;
        const int s = slots[q];
        if (threadIdx.x == 0) S = states[s];
        __syncthreads();
        if (S.status != kRunning) continue;
        // The attempts' copy of the descriptor: its pointers and sizes moved to
        // scalar registers, where they stay across the check calls (read through
        // the descriptor after a call, they would be vector loads, and every
        // address of the attempts would be formed in vector registers).
        LpDev L = lps[s];
        for (int h = 0; h < 2; ++h) {
            L.x[h] = uniform(L.x[h]);
            L.xa[h] = uniform(L.xa[h]);
            L.y[h] = uniform(L.y[h]);
            L.ya[h] = uniform(L.ya[h]);
            L.ax[h] = uniform(L.ax[h]);
        }
        L.c = uniform(L.c);
        L.lb = uniform(L.lb);
        L.ub = uniform(L.ub);
        L.rl = uniform(L.rl);
        L.ru = uniform(L.ru);
        L.tsi = uniform(L.tsi);
        L.tsv = uniform(L.tsv);
        L.asi = uniform(L.asi);
        L.asv = uniform(L.asv);
        L.tsv32 = uniform(L.tsv32);
        L.asv32 = uniform(L.asv32);
        L.lub = uniform(L.lub);
        L.rlu = uniform(L.rlu);
        L.n = __builtin_amdgcn_readfirstlane(L.n);
        L.m = __builtin_amdgcn_readfirstlane(L.m);
        double *ylds, *dxlds;
        int *tsblds, *asblds;
        unsigned char *tlen, *alen;
        lds_layout(dyn, L, ylds, dxlds, tsblds, asblds, tlen, alen);
        stage_layout<NT>(L, tsblds, asblds, tlen, alen);
        __syncthreads();
        bool ystaged = false;
        for (int period = 0;;) {
            if (S.phase != kAdvance) {
                check_period_call<NT>((const GLOBAL_AS LpDev *)(lps + s), (LDS_AS LpState *)&S, o,
                                      (LDS_AS double *)lds);
                ystaged = false;  // a restart may have replaced y
                if (S.status != kRunning) break;
            }
            if (threadIdx.x == 0)
                stop = period >= max_periods || (period >= min_periods && wall_clock64() - start >= budget);
            __syncthreads();
            if (stop) break;
            const long long remaining = S.iteration_limit - S.iterations;
            if (remaining <= 0) {
                __syncthreads();
                if (threadIdx.x == 0) S.status = kIterationLimit;
                break;
            }
            const long long count_ = remaining < o.check_every ? remaining : o.check_every;
            const double tau0 = S.step / S.weight, sigma0 = S.step * S.weight;
            const bool fp32 = !S.fp64;
            for (long long done = 0; done < count_;) {
                if (fp32)
                    attempt_lds<NT, 1, float>(L, S, o, tau0, sigma0, tiles, ylds, dxlds, tlen, alen, tsblds, asblds,
                                              powlds, &flag, ystaged);
                else
                    attempt_lds<NT, 1, double>(L, S, o, tau0, sigma0, tiles, ylds, dxlds, tlen, alen, tsblds,
                                               asblds, powlds, &flag, ystaged);
                if (flag == 1) ++done;
                else if (flag < 0) {
                    if (threadIdx.x == 0) S.status = kExhausted;
                    break;
                }
            }
            __syncthreads();
            if (threadIdx.x == 0 && S.status == kRunning) {
                S.iterations += count_;
                S.since_restart += count_;
                S.checks += 1;
                S.phase = kCheck;
                if (fp32) S.fp32_iterations += count_;
            }
            __syncthreads();
            if (S.status != kRunning) break;
            ++period;
        }
        __syncthreads();
        if (threadIdx.x == 0) states[s] = S;
    }
}

// ------------------------------------------------------------------ host --

using Clock = std::chrono::steady_clock;

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

// First-fit allocator over one device arena.
class Arena {
  public:
    void init(size_t bytes) {
        HIP_OK(hipMalloc(&base_, bytes));
        size_ = bytes;
        free_[0] = bytes;
    }
    ~Arena() {
        if (base_) hipFree(base_);
    }
    bool allocate(size_t bytes, size_t &offset) {
        bytes = align_up(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= bytes) {
                offset = it->first;
                size_t rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[offset + bytes] = rest;
                used_ += bytes;
                return true;
            }
        return false;
    }
    void release(size_t offset, size_t bytes) {
        bytes = align_up(bytes);
        used_ -= bytes;
        auto it = free_.emplace(offset, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    char *base() const { return static_cast<char *>(base_); }
    size_t size() const { return size_; }
    size_t used() const { return used_; }

  private:
    void *base_ = nullptr;
    size_t size_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;
};

struct Job {
    int64_t ticket = 0;
    std::vector<double> handoff_x, handoff_y;  // original-space warm start (current iterate) for the grid engine
    nyk::Model original;
    nyk::Prepared prep;
    std::unique_ptr<nyk::Verifier> verifier;
    nyk::Result result;
    bool finished = false;  // written by the scheduling thread only
    std::vector<double> x, y;  // scaled start point
    std::vector<double> warm_x, warm_y;  // original-space warm start as submitted (empty: cold)
    double step = 1, weight = 1;
    Clock::time_point start;
    int slot = -1;
    size_t offset = 0, bytes = 0;
    // Filled by a prepare worker before the job reaches the ready list.
    LpDev dev{};               // device pointers into the job's arena range
    int cls = 0;               // launch class
    double attempt_b = 0, attempt_b32 = 0, check_b = 0;
    bool terminal = false;     // the host decided the result without iterating
    bool resident = false;     // the image is in the arena at `offset`
    std::string prepare_error; // non-empty: prepare or upload failed
};

// One main stream plus one stream per launch class, created on first use and
// kept for the life of the process (default priority).
std::vector<hipStream_t> &stream_pool() {
    static std::vector<hipStream_t> pool = [] {
        std::vector<hipStream_t> p(2 + kClasses);
        for (auto &st : p) HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        return p;
    }();
    return pool;
}

// Splits CSR rows into consecutive blocks of at most kChunk nonzeros; a row
// longer than kChunk is a block of its own.  Returns block start rows plus
// the final row count.
std::vector<int64_t> row_blocks(const std::vector<int64_t> &ptr) {
    constexpr int64_t limit = kChunk;
    const int64_t rows = int64_t(ptr.size()) - 1;
    std::vector<int64_t> b{0};
    for (int64_t r = 0; r < rows;) {
        const int64_t start = r;
        if (ptr[r + 1] - ptr[r] > kChunk)
            ++r;
        else
            while (r < rows && ptr[r + 1] - ptr[r] <= kChunk && ptr[r + 1] - ptr[start] <= limit) ++r;
        b.push_back(r);
    }
    return b;
}

// The rows of a CSR matrix longer than kChunk, in order (the checks sum them
// with the workgroup's tree).
std::vector<int64_t> long_rows(const std::vector<int64_t> &ptr) {
    std::vector<int64_t> r;
    for (size_t i = 0; i + 1 < ptr.size(); ++i)
        if (ptr[i + 1] - ptr[i] > kChunk) r.push_back(int64_t(i));
    return r;
}

// Offsets of the transpose of a CSR matrix (column counts, prefix-summed).
std::vector<int64_t> t_ptr(const nyk::Model &m) {
    std::vector<int64_t> p(m.cols() + 1, 0);
    for (auto j : m.A.index) p[j + 1]++;
    for (int64_t j = 0; j < m.cols(); ++j) p[j + 1] += p[j];
    return p;
}

// A CSR matrix in 64-row slices (rows_dot_sell): slice s holds its rows'
// entries entry-major from base[s], padded to the slice's longest row.
struct Sliced {
    std::vector<int64_t> base, index;
    std::vector<double> value;
    std::vector<float> value32;  // with `fp32`: the values at their indices' positions
};

// The sliced layout of rows_dot_sell: per slice of 64 rows an even width w
// (the longest row, rounded up); the value of entry u of row r at
// base + (u / 2) * 128 + (r % 64) * 2 + u % 2, its index at
// base + (u / 4) * 256 + (r % 64) * 4 + u % 4, or, in the last pair of a
// width that is not a multiple of 4, at base + (u / 4) * 256 + (r % 64) * 2 + u % 2.
Sliced slice_rows(const std::vector<int64_t> &ptr, const std::vector<int64_t> &index, const std::vector<double> &value,
                  bool fp32 = false) {
    const int64_t rows = int64_t(ptr.size()) - 1, slices = (rows + 63) / 64;
    Sliced z;
    z.base.assign(slices + 1, 0);
    for (int64_t sl = 0; sl < slices; ++sl) {
        int64_t width = 0;
        for (int64_t r = sl * 64; r < std::min(rows, sl * 64 + 64); ++r) width = std::max(width, ptr[r + 1] - ptr[r]);
        z.base[sl + 1] = z.base[sl] + (width + 1) / 2 * 2 * 64;
    }
    z.index.assign(z.base[slices], 0);
    z.value.assign(z.base[slices], 0.0);
    if (fp32) z.value32.assign(z.base[slices], 0.f);
    for (int64_t r = 0; r < rows; ++r)
        for (int64_t u = 0; u < ptr[r + 1] - ptr[r]; ++u) {
            const int64_t b = z.base[r / 64], w = (z.base[r / 64 + 1] - b) / 64;
            const int64_t quad = u / 4 * 4 + 4 <= w;
            const int64_t at = b + u / 4 * 256 + (quad ? r % 64 * 4 + u % 4 : r % 64 * 2 + u % 2);
            z.value[b + u / 2 * 128 + r % 64 * 2 + u % 2] = value[ptr[r] + u];
            z.index[at] = index[ptr[r] + u];
            if (fp32) z.value32[at] = float(value[ptr[r] + u]);
        }
    return z;
}

// Builds one LP's device image: every array at a 256-byte aligned offset.
struct Layout {
    size_t total = 0;
    template <class T>
    size_t add(size_t count) {
        size_t at = total;
        total = align_up(total + std::max<size_t>(count, 1) * sizeof(T));
        return at;
    }
};

// Launch shapes of a crowded LDS class's step kernel: threads per workgroup
// and rows per thread.  Every shape gives the same results (tile-ordered
// sums), so the choice is free to follow measurements.
struct Shape {
    int threads, rows;
};
constexpr Shape kShapes[] = {{256, 1}, {256, 2}, {256, 4}, {1024, 1}};
constexpr int kShapeCount = int(sizeof kShapes / sizeof kShapes[0]);
static_assert(kShapeCount == 4, "niryukti_batch_stats holds four shape entries");

// Autotuner of the crowded LDS classes' launch shapes, one class at a time
// (coordinate descent from the start shapes, which were measured best on
// gfx950): each shape of the class under trial runs kRepeats scheduling
// rounds, interleaved with the other shapes in mirrored order (ABCD DCBA
// ABCD DCBA: a drift of the rounds' cost falls on every shape alike; the
// rounds still do full work), while the other classes keep their current
// shapes.  The class then keeps the shape with the highest mean round
// throughput -- algorithmic bytes of every crowded launch of the round over
// the round's device time -- when it beats the current shape by at least
// kMargin (single rounds of one class differ by several percent with no
// change of shape, and in streamed runs the means of the trials still differ
// by up to about 10 %: with an 8 % margin the smallest class moved to 256 x 2
// in some runs, which measured 4-8 % slower end to end).  The score is the round's and not the class's own: the
// classes run side by side, and a shape with more workgroups per CU takes
// slots from its neighbours (scored alone, such a shape won its class's
// trials and lost end to end).  Trials count only once the class under trial
// has run enough LPs to fill the device with its start shape -- two
// 256-thread workgroups per CU of the 256 for the smallest class, one
// 1,024-thread workgroup per CU for the larger two -- for kSettle rounds:
// below that a launch lasts as long as its slowest LP, whatever the shape,
// and the first rounds of a batch run while prepares still admit LPs.  (A
// fixed 512 left the larger classes untuned in streamed runs, where fewer of
// their LPs are resident at once.)
// The smallest class (m + n <= 2,048) tries the 256-thread shapes only: a
// 1,024-thread workgroup has at most two rows or columns per thread there and
// idles most of its waves in each pass (its trials were within the rounds'
// noise and moved the class in some runs and not in others).
class ShapeTuner {
  public:
    static constexpr int kTuned = 3;  // the crowded LDS classes
    static constexpr int kMinTrialLps[kTuned] = {512, 256, 256};  // NIRYUKTI_BATCH_TUNE_MIN_LPS overrides them
    static constexpr int kSettle = 2;
    static constexpr int kRepeats = 4;
    static constexpr double kMargin = 0.15;
    static bool allowed(int c, int shape) { return c > 0 || kShapes[shape].threads <= 256; }
    void start_with(int c, int shape) { current_[c] = shape; }
    void set_min_lps(int lps) {
        for (int &v : min_lps_) v = std::max(1, lps);
    }
    void fix(int c, int shape) {
        current_[c] = shape;
        fixed_[c] = true;
    }
    bool decided(int c) const { return fixed_[c] || done_[c]; }
    int current(int c) const { return current_[c]; }
    // The shape class c runs in the next round.
    int next(int c) const { return c == trial_class() && settled_ >= kSettle ? trial_shape(tried_) : current_[c]; }
    // One round: `bytes` of every crowded launch over `seconds` of device time.
    void record(const int *shape_of, const int *lps, double bytes, double seconds) {
        const int c = trial_class();
        if (c < 0) return;
        if (lps[c] < min_lps_[c]) {
            settled_ = 0;  // trials so far stand; they resume after kSettle full rounds
            return;
        }
        if (settled_ < kSettle) {
            ++settled_;
            return;
        }
        if (shape_of[c] != trial_shape(tried_) || !(bytes > 0) || !(seconds > 0)) return;
        rate_[shape_of[c]] += bytes / seconds;
        if (++tried_ < kRepeats * candidates(c)) return;
        int best = current_[c];
        for (int k = 0; k < kShapeCount; ++k)
            if (allowed(c, k) && rate_[k] > rate_[best]) best = k;
        if (rate_[best] >= (1 + kMargin) * rate_[current_[c]]) current_[c] = best;
        done_[c] = true;
        tried_ = 0;
        for (double &r : rate_) r = 0;
    }

  private:
    int trial_class() const {
        for (int c = 0; c < kTuned; ++c)
            if (!fixed_[c] && !done_[c]) return c;
        return -1;
    }
    static int candidates(int c) {
        int k = 0;
        for (int s = 0; s < kShapeCount; ++s) k += allowed(c, s);
        return k;
    }
    // Trial t of the class under trial: its current shape first, then the
    // other candidates by index, mirrored every other sweep.
    int trial_shape(int t) const {
        const int c = trial_class();
        int order[kShapeCount], k = 0;
        order[k++] = current_[c];
        for (int s = 0; s < kShapeCount; ++s)
            if (s != current_[c] && allowed(c, s)) order[k++] = s;
        const int i = t % k;
        return (t / k) % 2 ? order[k - 1 - i] : order[i];
    }
    int current_[kTuned] = {};
    bool fixed_[kTuned] = {}, done_[kTuned] = {};
    int tried_ = 0, settled_ = 0, min_lps_[kTuned] = {kMinTrialLps[0], kMinTrialLps[1], kMinTrialLps[2]};
    double rate_[kShapeCount] = {};
};

class BatchEngine {
  public:
    BatchEngine(const nyk::Options &o, int64_t arena_bytes, int64_t max_slots) : opt_(o) {
        if (o.method != "pdhg") throw std::runtime_error("The batched engine runs the pdhg method");
        mixed_ = o.matrix_precision == "mixed";
        if (arena_bytes <= 0 || max_slots <= 0) throw std::runtime_error("Invalid arena size or slot count");
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count < 1) throw std::runtime_error("No HIP device is visible");
        max_slots_ = int(max_slots);
        arena_.init(size_t(arena_bytes));
        // Streams come from a per-process pool created once, so repeated
        // engines keep the same stream-to-hardware-queue mapping.
        auto &pool = stream_pool();
        stream_ = pool[0];
        for (int c = 0; c <= kClasses; ++c) class_stream_[c] = pool[1 + c];
        // Shapes measured best on gfx950 (tried first): 256 threads x 1 row
        // up to m + n = 2,048, 1,024 x 1 above.  NIRYUKTI_BATCH_ROWS_PER_THREAD
        // (1, 2 or 4) fixes the smallest class at 256 threads x that many rows;
        // NIRYUKTI_BATCH_TUNE=0 keeps the measured shapes without trials;
        // NIRYUKTI_BATCH_TUNE_MIN_LPS=n lets rounds with n LPs in the class
        // count as trials (a small batch then reaches a decision);
        // NIRYUKTI_BATCH_SHAPES=a,b,c fixes the three classes' shapes (indices
        // into kShapes), e.g. so that profiler runs of one workload launch the
        // same kernels.
        static_assert(kClasses - 1 == ShapeTuner::kTuned, "one tuned shape per crowded LDS class");
        for (int c = 0; c < kClasses - 1; ++c) tuner_.start_with(c, c == 0 ? 0 : 3);
        if (const char *env = std::getenv("NIRYUKTI_BATCH_ROWS_PER_THREAD")) {
            const int r = std::atoi(env);
            tuner_.fix(0, r == 4 ? 2 : r == 2 ? 1 : 0);
        }
        if (const char *env = std::getenv("NIRYUKTI_BATCH_TUNE"))
            if (std::atoi(env) == 0)
                for (int c = 0; c < kClasses - 1; ++c) tuner_.fix(c, tuner_.current(c));
        if (const char *env = std::getenv("NIRYUKTI_BATCH_TUNE_MIN_LPS")) tuner_.set_min_lps(std::atoi(env));
        if (const char *env = std::getenv("NIRYUKTI_BATCH_SHAPES")) {
            int shape[kClasses - 1];
            if (std::sscanf(env, "%d,%d,%d", &shape[0], &shape[1], &shape[2]) != kClasses - 1)
                throw std::runtime_error("NIRYUKTI_BATCH_SHAPES takes three shape indices, e.g. 0,3,3");
            for (int c = 0; c < kClasses - 1; ++c) {
                if (shape[c] < 0 || shape[c] >= kShapeCount)
                    throw std::runtime_error("NIRYUKTI_BATCH_SHAPES: shape index out of range");
                tuner_.fix(c, shape[c]);
            }
        }
        // Class capacities (m + n) bounded by what one workgroup may allocate
        // next to the kernel's static LDS.
        HIP_OK(hipGetDevice(&device_));
        const int device = device_;
        hipDeviceProp_t prop;
        HIP_OK(hipGetDeviceProperties(&prop, device));
        const long long room = (long long)prop.sharedMemPerBlock - kStaticLds - 16;
        int wall_khz = 0;
        HIP_OK(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device));
        if (wall_khz <= 0) throw std::runtime_error("The device reports no constant-rate clock");
        tail_budget_ticks_ = (long long)wall_khz * kTailBudgetMs;
        compute_units_ = prop.multiProcessorCount;
        for (int c = 0; c < kClasses - 1; ++c) {
            int cap = kClassDoubles[c];
            while (cap > 0 && (long long)lds_bytes(cap) > room) --cap;
            class_cap_[c] = cap;
        }
        HIP_OK(hipMalloc(&d_desc_, sizeof(LpDev) * max_slots_));
        HIP_OK(hipMalloc(&d_state_, sizeof(LpState) * max_slots_));
        HIP_OK(hipMalloc(&d_slots_, sizeof(int) * max_slots_ * (kClasses + 1)));
        HIP_OK(hipMalloc(&d_queue_, sizeof(unsigned)));
        // k^-0.3 and k^-0.6 by std::pow for every k the step rule can reach
        // (k = accepted + 2 <= iteration_limit + 1), up to 2^22 entries.
        pow_count_ = std::min<long long>(std::max<long long>(opt_.iteration_limit, 0) + 3, 1LL << 22);
        {
            std::vector<double> p3(static_cast<size_t>(pow_count_)), p6(static_cast<size_t>(pow_count_));
            for (long long k = 0; k < pow_count_; ++k) {
                p3[size_t(k)] = std::pow(double(k), -.3);
                p6[size_t(k)] = std::pow(double(k), -.6);
            }
            HIP_OK(hipMalloc(&d_pow3_, sizeof(double) * size_t(pow_count_)));
            HIP_OK(hipMalloc(&d_pow6_, sizeof(double) * size_t(pow_count_)));
            HIP_OK(hipMemcpyAsync(d_pow3_, p3.data(), sizeof(double) * p3.size(), hipMemcpyHostToDevice, stream_));
            HIP_OK(hipMemcpyAsync(d_pow6_, p6.data(), sizeof(double) * p6.size(), hipMemcpyHostToDevice, stream_));
            HIP_OK(hipStreamSynchronize(stream_));
        }
        HIP_OK(hipEventCreate(&uploaded_));

        desc_.resize(max_slots_);
        state_.resize(max_slots_);
        slot_job_.assign(max_slots_, -1);
        slot_bytes_.assign(max_slots_, 0);
        slot_bytes32_.assign(max_slots_, 0);
        slot_check_bytes_.assign(max_slots_, 0);
        slot_class_.assign(max_slots_, kClasses - 1);
        slot_group_.assign(max_slots_, kClasses - 1);
        for (int s = max_slots_ - 1; s >= 0; --s) free_slots_.push_back(s);
    }
    ~BatchEngine() {
        if (settling_) {
            try {
                settling_->task.wait();
            } catch (...) {
            }
        }
        {
            std::lock_guard<std::mutex> lock(handoff_mu_);
            stopping_ = true;
            handoff_queue_.clear();
        }
        handoff_cv_.notify_all();
        for (auto &w : workers_) w.join();
        {
            std::lock_guard<std::mutex> lock(prep_mu_);
            prep_stop_ = true;
            prep_todo_.clear();
        }
        prep_cv_.notify_all();
        {
            std::lock_guard<std::mutex> lock(arena_mu_);
        }
        arena_cv_.notify_all();
        for (auto &w : preparers_) w.join();
        // a round may still be in flight on the class streams
        if (stream_) hipStreamSynchronize(stream_);
        for (auto &cs : class_stream_)
            if (cs) hipStreamSynchronize(cs);
        hipFree(d_desc_);
        hipFree(d_state_);
        hipFree(d_slots_);
        hipFree(d_queue_);
        hipFree(d_pow3_);
        hipFree(d_pow6_);
        if (d_stage_) hipFree(d_stage_);
        if (h_stage_) hipHostFree(h_stage_);
        if (d_pieces_) hipFree(d_pieces_);
        for (auto &list : class_events_)
            for (auto &e : list) hipEventDestroy(e);
        if (uploaded_) hipEventDestroy(uploaded_);
    }

    // Any thread, also while the scheduling thread is inside step(): the copy
    // and the checks run on the caller's thread; the job reaches jobs_ when
    // the scheduling thread next drains incoming_.
    int64_t submit(const niryukti_lp &lp, const double *x0, const double *y0) {
        auto job = std::make_unique<Job>();
        job->start = Clock::now();
        nyk::Model &m0 = job->original;
        const int64_t nnz = lp.ptr[lp.rows];
        m0.A.rows = lp.rows;
        m0.A.cols = lp.cols;
        m0.A.ptr.assign(lp.ptr, lp.ptr + lp.rows + 1);
        m0.A.index.assign(lp.index, lp.index + nnz);
        m0.A.value.assign(lp.value, lp.value + nnz);
        m0.c.assign(lp.c, lp.c + lp.cols);
        m0.lb.assign(lp.lb, lp.lb + lp.cols);
        m0.ub.assign(lp.ub, lp.ub + lp.cols);
        m0.rl.assign(lp.rl, lp.rl + lp.rows);
        m0.ru.assign(lp.ru, lp.ru + lp.rows);
        m0.offset = lp.offset;
        m0.sense = lp.sense == 0 ? 1 : lp.sense;
        m0.validate();
        if (m0.rows() >= (1LL << 31) || m0.cols() >= (1LL << 31) || nnz >= (1LL << 31))
            throw std::runtime_error("Batched LPs must have fewer than 2^31 rows, columns and nonzeros");
        if (x0) job->warm_x.assign(x0, x0 + lp.cols);
        if (y0) job->warm_y.assign(y0, y0 + lp.rows);
        for (double v : job->warm_x)
            if (!std::isfinite(v)) throw std::runtime_error("Nonfinite warm start");
        for (double v : job->warm_y)
            if (!std::isfinite(v)) throw std::runtime_error("Nonfinite warm start");
        Job *j = job.get();
        {
            std::lock_guard<std::mutex> lock(submit_mu_);
            start_preparers();
        }
        // the ticket is the job's sequence number in incoming_, so a drain
        // appends each job to jobs_ at its own index
        const int64_t ticket = incoming_.push(std::move(job), [](std::unique_ptr<Job> &q, int64_t n) { q->ticket = n; });
        ++prep_outstanding_;
        {
            std::lock_guard<std::mutex> lock(prep_mu_);
            prep_todo_.push_back(j);
        }
        prep_cv_.notify_one();
        return ticket;
    }

    // Scheduling thread: moves submitted jobs into jobs_ (indexed by ticket).
    void drain_incoming() { incoming_.drain(jobs_); }

    // ---- host prepare off the launch path ----
    //
    // Presolve, scaling, the start-point checks, the device image (sliced
    // copies, row blocks, transposes) and its upload into the arena run on
    // prepare workers, so the scheduling thread only assigns slots and
    // launches.  A worker that finds no arena room waits until a retired LP
    // frees some.  Results do not depend on when an LP is admitted.
    void start_preparers() {
        if (!preparers_.empty()) return;
        const int device = device_;
        const int workers = std::max(1, nyk::host_threads());
        for (int w = 0; w < workers; ++w) preparers_.emplace_back([this, device] { prepare_loop(device); });
    }

    void prepare_loop(int device) {
        // Prepares are background work beside the scheduling thread, which
        // drives the device: with one worker per core they would otherwise
        // delay it when a streamed pass arrives (the device then idles between
        // rounds).  A lower priority for the workers (a higher nice value,
        // which needs no privilege; ignored where the call fails) cut that gap.
        (void)setpriority(PRIO_PROCESS, id_t(syscall(SYS_gettid)), kPrepareNice);
        (void)hipSetDevice(device);
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) st = nullptr;
        for (;;) {
            Job *j = nullptr;
            {
                std::unique_lock<std::mutex> lock(prep_mu_);
                prep_cv_.wait(lock, [&] { return prep_stop_ || !prep_todo_.empty(); });
                if (prep_stop_) break;
                j = prep_todo_.front();
                prep_todo_.pop_front();
            }
            try {
                if (!st) throw std::runtime_error("Could not create a prepare stream");
                prepare_job(*j, st);
            } catch (const std::exception &e) {
                j->prepare_error = e.what();
                if (j->prepare_error.empty()) j->prepare_error = "unknown error";
            }
            {
                std::lock_guard<std::mutex> lock(prep_mu_);
                prep_ready_.push_back(j);
            }
            ready_cv_.notify_all();
        }
        if (st) (void)hipStreamDestroy(st);
    }

    // setup() + the device image + its upload, on a prepare worker.
    void prepare_job(Job &j, hipStream_t st) {
        nyk::Options o = opt_;
        o.initial_x = j.warm_x;
        o.initial_y = j.warm_y;
        j.terminal = setup(j, o);
        if (j.terminal) return;
        j.cls = lds_class(j);
        j.attempt_b = attempt_bytes(j);
        j.attempt_b32 = attempt_bytes(j, true);
        j.check_b = check_bytes(j);
        std::vector<char> image;
        const size_t bytes = build_image(j, image, j.dev, [&](size_t total) {
            j.offset = reserve_arena(total);
            j.bytes = total;
            j.resident = true;
            return arena_.base() + j.offset;
        });
        HIP_OK(hipMemcpyAsync(arena_.base() + j.offset, image.data(), bytes, hipMemcpyHostToDevice, st));
        HIP_OK(hipStreamSynchronize(st));
    }

    // First-fit room for `bytes` in the arena; waits while resident LPs hold
    // it.  An LP larger than the whole arena can never be admitted.
    size_t reserve_arena(size_t bytes) {
        std::unique_lock<std::mutex> lock(arena_mu_);
        if (align_up(bytes) > arena_.size())
            throw std::runtime_error("LP does not fit the batch arena (" + std::to_string(bytes) + " bytes)");
        size_t offset = 0;
        while (!arena_.allocate(bytes, offset)) {
            if (prep_stop_) throw std::runtime_error("Batch is closing");
            arena_cv_.wait(lock);
        }
        return offset;
    }
    void release_arena(size_t offset, size_t bytes) {
        {
            std::lock_guard<std::mutex> lock(arena_mu_);
            arena_.release(offset, bytes);
        }
        arena_cv_.notify_all();
    }

    // Moves prepared jobs to the admission queue or, when the host already
    // decided them, to the finished list (scheduling thread only).  With
    // `wait`, blocks until at least one prepare has finished.
    void collect_prepared(bool wait) {
        std::vector<Job *> ready;
        {
            std::unique_lock<std::mutex> lock(prep_mu_);
            if (wait && prep_outstanding_ > 0)
                ready_cv_.wait(lock, [&] { return !prep_ready_.empty(); });
            ready.swap(prep_ready_);
        }
        // every job taken from prep_ready_ entered incoming_ before its
        // prepare started (submit_mu_, then prep_mu_), so it is drained here
        drain_incoming();
        std::sort(ready.begin(), ready.end(), [](const Job *a, const Job *b) { return a->ticket < b->ticket; });
        for (Job *j : ready) {
            --prep_outstanding_;
            if (!j->prepare_error.empty()) {
                if (j->resident) release_arena(j->offset, j->bytes);
                j->resident = false;
                const bool room = j->prepare_error.rfind("LP does not fit", 0) == 0;
                j->result.message = room ? j->prepare_error : "Prepare failed: " + j->prepare_error;
                retire_host(*j, room ? "UNSUPPORTED" : "NUMERICAL_ERROR");
            } else if (j->terminal) {
                j->finished = true;
                done_.push_back(j->ticket);
            } else {
                queue_.push_back(j->ticket);
            }
        }
    }

    // One call: finish the round launched by the previous call (wait for
    // it, read back every LP's state, account its time and bytes, stop LPs
    // past their time limit), admit queued LPs into free slots, launch the
    // next round, and only then settle the finished round's stopped LPs
    // (certification, retirement, resumption) while the new round runs.  The
    // device so idles between two rounds only for the read-back, admission
    // and launch, not for certification or the caller's own work between
    // calls.  An LP that stops in one round is left out of the next (its
    // stop is known before the next round is planned); its slot and arena
    // space are freed after that round is launched, so a new LP takes them
    // one round later.
    int64_t step(int checks) {
        drain_incoming();
        std::vector<int> finished;
        if (round_) finished = finish_round();
        const auto ta = Clock::now();
        admit();
        admit_seconds_ += std::chrono::duration<double>(Clock::now() - ta).count();
        launch_round(checks);
        settle(finished);
        drain_handoffs(!round_);
        int64_t resident = 0;
        for (int s = 0; s < max_slots_; ++s) resident += slot_job_[s] >= 0;
        // Nothing to launch yet: wait for a prepare worker instead of spinning.
        if (!round_ && !resident && queue_.empty() && prep_outstanding_ > 0) collect_prepared(true);
        return int64_t(queue_.size()) + handoffs_outstanding_ + prep_outstanding_ + resident + (round_ ? 1 : 0);
    }

    // A launched scheduling round: its launch groups, the LPs it runs and
    // their counters at launch (for the round's statistics).
    struct Round {
        struct Group {
            int kind;  // 0..kClasses-1: that class; kClasses: the tail group
            int cap;
            std::vector<int> slots;
            int fp32 = 0;  // mixed precision: the first `fp32` LPs step with FP32 values
        };
        std::vector<Group> groups;
        std::vector<int> active;
        std::vector<long long> before, checks_before;
        std::vector<char> fp32;  // the LP stepped with FP32 values at launch
        int shape_of[kClasses - 1] = {};
        int checks = 0;
    };

    // Plans a round from the host's view of every LP and launches it without
    // waiting (no round is in flight: the previous one was finished first, so
    // the whole state array can be uploaded).
    void launch_round(int checks) {
        auto R = std::make_unique<Round>();
        // mixed precision: a crowded LDS class lists its LPs on FP32 values
        // first and those on FP64 values after them (by_fp64); each check
        // period launches the step kernel in FP32 values over the first part,
        // in FP64 values over the second, and one check over both
        std::vector<int> by_class[kClasses], by_fp64[kClasses];
        for (int s = 0; s < max_slots_; ++s)
            if (slot_job_[s] >= 0 && state_[s].status == kRunning) {
                R->active.push_back(s);
                if (mixed_ && state_[s].fp64 && slot_class_[s] < kClasses - 1) by_fp64[slot_class_[s]].push_back(s);
                else by_class[slot_class_[s]].push_back(s);
            }
        int split[kClasses];
        for (int c = 0; c < kClasses; ++c) {
            split[c] = int(by_class[c].size());
            by_class[c].insert(by_class[c].end(), by_fp64[c].begin(), by_fp64[c].end());
        }
        if (R->active.empty()) return;
        R->checks = checks;
        if (dirty_) {
            HIP_OK(hipMemcpyAsync(d_desc_, desc_.data(), sizeof(LpDev) * max_slots_, hipMemcpyHostToDevice, stream_));
            dirty_ = false;
        }
        // from a copy: settle changes state_ (of LPs outside this round)
        // while the upload may still be reading it
        upload_ = state_;
        HIP_OK(hipMemcpyAsync(d_state_, upload_.data(), sizeof(LpState) * max_slots_, hipMemcpyHostToDevice, stream_));
        // Launch groups: each crowded LDS class, the global class, and one
        // tail group that gathers every LDS class with few LPs left, so a
        // handful of slow LPs share one launch per check period instead
        // of one per class (launches of different classes can land on the
        // same hardware queue and then run one after another).
        auto &groups = R->groups;
        Round::Group tail{kClasses, 0, {}};
        for (int c = 0; c < kClasses; ++c) {
            if (by_class[c].empty()) continue;
            if (c < kClasses - 1 && int(by_class[c].size()) <= kTailLps) {
                tail.cap = std::max(tail.cap, class_cap_[c]);
                tail.slots.insert(tail.slots.end(), by_class[c].begin(), by_class[c].end());
            } else {
                groups.push_back(Round::Group{c, c < kClasses - 1 ? class_cap_[c] : 0, by_class[c], split[c]});
            }
        }
        const size_t tail_lps = tail.slots.size();
        if (!tail.slots.empty()) groups.push_back(std::move(tail));
        for (const auto &g : groups)
            HIP_OK(hipMemcpyAsync(d_slots_ + size_t(g.kind) * max_slots_, g.slots.data(), sizeof(int) * g.slots.size(),
                                  hipMemcpyHostToDevice, stream_));
        if (tail_lps) HIP_OK(hipMemcpyAsync(d_queue_, &zero_, sizeof(unsigned), hipMemcpyHostToDevice, stream_));
        HIP_OK(hipEventRecord(uploaded_, stream_));
        BatchOpts bo{opt_.tol, opt_.check_every, opt_.adaptive ? 1 : 0, opt_.restart ? 1 : 0, mixed_ ? 1 : 0,
                     handoff_,    d_pow3_,         d_pow6_,                  pow_count_};
        R->before.resize(R->active.size());
        R->checks_before.resize(R->active.size());
        R->fp32.resize(R->active.size());
        for (size_t k = 0; k < R->active.size(); ++k) {
            R->before[k] = state_[R->active[k]].accepted + state_[R->active[k]].rejected;
            R->checks_before[k] = state_[R->active[k]].checks;
            R->fp32[k] = !state_[R->active[k]].fp64;
        }
        for (int c = 0; c < kClasses - 1; ++c) R->shape_of[c] = tuner_.next(c);
        // The tail launch continues past `checks` periods only when it
        // has the round to itself and no LP waits for admission: beside
        // crowded launches (which take `checks` periods) a longer tail
        // launch would hold the round, and with it every crowded LP,
        // until its budget ran out; with LPs prepared or in preparation
        // it would hold their admission (between two streamed passes, while
        // the next pass is prepared, the tail group holds hundreds of LPs).
        bool crowded = false;
        for (const auto &g : groups) crowded = crowded || g.kind != kClasses;
        const bool waiting = !queue_.empty() || prep_outstanding_ > 0;
        const long long tail_budget = crowded || waiting ? 0 : tail_budget_ticks_;
        for (const auto &g : groups) {
            auto &ev = class_events_[g.kind];
            while (ev.size() < size_t(2 * checks + 2)) {
                hipEvent_t e;
                HIP_OK(hipEventCreate(&e));
                ev.push_back(e);
            }
            hipStream_t cs = class_stream_[g.kind];
            HIP_OK(hipStreamWaitEvent(cs, uploaded_, 0));
            const int grid = int(g.slots.size());
            const int *slots = d_slots_ + size_t(g.kind) * max_slots_;
            HIP_OK(hipEventRecord(ev[0], cs));
            if (g.kind == kClasses) {
                launch_tail(g.cap, grid, cs, slots, bo, checks, tail_budget);
                HIP_OK(hipEventRecord(ev[1], cs));
            } else {
                for (int k = 0; k < checks; ++k) {
                    if (g.kind < kClasses - 1 && mixed_) {
                        if (g.fp32) launch_lds(g.cap, R->shape_of[g.kind], g.fp32, cs, slots, bo, true);
                        if (grid > g.fp32)
                            launch_lds(g.cap, R->shape_of[g.kind], grid - g.fp32, cs, slots + g.fp32, bo, false);
                    } else if (g.kind < kClasses - 1)
                        launch_lds(g.cap, R->shape_of[g.kind], grid, cs, slots, bo, false);
                    else
                        batch_advance_kernel<<<grid, kThreads, 0, cs>>>(d_desc_, d_state_, slots, bo);
                    HIP_OK(hipEventRecord(ev[2 * k + 1], cs));
                    batch_check_kernel<<<grid, kThreads, 0, cs>>>(d_desc_, d_state_, slots, bo);
                    HIP_OK(hipEventRecord(ev[2 * k + 2], cs));
                }
            }
            HIP_OK(hipGetLastError());
        }
        round_ = std::move(R);
    }

    // Waits for the round in flight, reads back the states of its LPs and
    // accounts the round; stops running LPs past their time limit (they are
    // left out of the next round and retired by settle).  Returns the
    // round's LPs.  Only the round's own slots take the device's values: the
    // host may have changed other slots meanwhile (an LP resumed after a
    // failed certification, a newly admitted one), and the next launch
    // uploads those changes.
    std::vector<int> finish_round() {
        std::unique_ptr<Round> R = std::move(round_);
        // stream_ waits for the round only here, so that settle's downloads
        // on stream_ run beside the round
        for (const auto &g : R->groups)
            HIP_OK(hipStreamWaitEvent(stream_, class_events_[g.kind][g.kind == kClasses ? 1 : 2 * R->checks], 0));
        readback_.resize(size_t(max_slots_));
        HIP_OK(hipMemcpyAsync(readback_.data(), d_state_, sizeof(LpState) * max_slots_, hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
        for (int s : R->active) state_[s] = readback_[s];
        const int checks = R->checks;
        float span = 0;
        const auto &groups = R->groups;
        for (const auto &g : groups) {
            auto &ev = class_events_[g.kind];
            if (g.kind == kClasses) {
                // the fused launch: attempts and checks together
                float a = 0;
                HIP_OK(hipEventElapsedTime(&a, ev[0], ev[1]));
                advance_seconds_ += a * 1e-3;
                class_seconds_[g.kind] += a * 1e-3;
                class_launches_[g.kind] += 1;
            }
            for (int k = 0; k < (g.kind == kClasses ? 0 : checks); ++k) {
                float a = 0, b = 0;
                HIP_OK(hipEventElapsedTime(&a, ev[2 * k], ev[2 * k + 1]));
                HIP_OK(hipEventElapsedTime(&b, ev[2 * k + 1], ev[2 * k + 2]));
                advance_seconds_ += a * 1e-3;
                check_seconds_ += b * 1e-3;
                class_seconds_[g.kind] += a * 1e-3;
                class_launches_[g.kind] += 1;
            }
            float t = 0;  // from the round's upload to the group's last event
            HIP_OK(hipEventElapsedTime(&t, uploaded_, ev[g.kind == kClasses ? 1 : 2 * checks]));
            span = std::max(span, t);
            for (int s : g.slots) slot_group_[s] = g.kind;
        }
        double group_seconds[kClasses + 1] = {}, group_bytes[kClasses + 1] = {};
        int group_lps[kClasses + 1] = {};
        for (const auto &g : groups) group_lps[g.kind] = int(g.slots.size());
        for (const auto &g : groups) {
            auto &ev = class_events_[g.kind];
            for (int k = 0; k < (g.kind == kClasses ? 0 : checks); ++k) {
                float a = 0;
                HIP_OK(hipEventElapsedTime(&a, ev[2 * k], ev[2 * k + 1]));
                group_seconds[g.kind] += a * 1e-3;
            }
        }
        const auto &active = R->active;
        for (size_t k = 0; k < active.size(); ++k) {
            const LpState &st = state_[active[k]];
            const long long done = st.accepted + st.rejected - R->before[k];
            const double bytes = R->fp32[k] ? slot_bytes32_[active[k]] : slot_bytes_[active[k]];
            group_bytes[slot_group_[active[k]]] += double(done) * bytes;
            attempts_ += done;
            advance_bytes_ += double(done) * bytes;
            class_bytes_[slot_group_[active[k]]] += double(done) * bytes;
            if (slot_group_[active[k]] == kClasses)  // the fused tail launch also did the checks
                class_bytes_[kClasses] += double(st.checks - R->checks_before[k]) * slot_check_bytes_[active[k]];
        }
        const double round_seconds = span * 1e-3;
        tuner_.record(R->shape_of, group_lps, group_bytes[0] + group_bytes[1] + group_bytes[2], round_seconds);
        for (int c = 0; c < kClasses - 1; ++c)
            if (group_lps[c]) {
                shape_seconds_[R->shape_of[c]] += group_seconds[c];
                shape_bytes_[R->shape_of[c]] += group_bytes[c];
                shape_launches_[R->shape_of[c]] += checks;
            }
        device_seconds_ += round_seconds;
        ++launches_;
        // LPs past their time limit stop here (settle retires them with
        // TIME_LIMIT), before the next round is planned.
        for (int s : active) {
            Job &j = *jobs_[slot_job_[s]];
            if (state_[s].status == kRunning &&
                std::chrono::duration<double>(Clock::now() - j.start).count() >= opt_.time_limit) {
                state_[s].status = kIterationLimit;
                j.result.status = "TIME_LIMIT";
            }
        }
        return active;
    }

    // Handoff of long-running LPs to the grid-wide single-LP engine: after
    // `iterations` iterations in the batch an LP continues on its own, warm
    // started from its current iterate, on `workers` host threads
    // (each solve has its own HIP stream).  One workgroup walks a small LP
    // through a chain of dependent memory round trips per step; the grid
    // engine spreads the same step over the whole GPU, so the slow tail of a
    // batch finishes several times sooner.  0 iterations disables it.
    void set_handoff(int64_t iterations, int workers) {
        if (iterations < 0 || workers < 1 || workers > 16) throw std::runtime_error("Invalid handoff settings");
        if (handoffs_outstanding_) throw std::runtime_error("Handoff settings cannot change while LPs are handed off");
        handoff_ = iterations;
        handoff_workers_ = workers;
    }

    // Tickets finished since the last call, in finishing order.
    int64_t finished(int64_t *tickets, int64_t capacity) {
        int64_t k = 0;
        while (k < capacity && !done_.empty()) {
            tickets[k++] = done_.front();
            done_.pop_front();
        }
        return k;
    }

    // Finished tickets with their results in one call (x and y are not
    // copied), each released after its copy when `release_after`.
    int64_t collect(int64_t *tickets, niryukti_result *out, int64_t capacity, bool release_after) {
        drain_incoming();
        int64_t k = 0;
        while (k < capacity && !done_.empty()) {
            const int64_t t = done_.front();
            done_.pop_front();
            std::memset(&out[k], 0, sizeof out[k]);  // x = y = NULL, strings padded with zeros
            result(t, &out[k]);
            tickets[k++] = t;
            if (release_after) release(t);
        }
        return k;
    }

    // Frees the host copy of a finished LP (model, prepare maps, result).
    // The memory is returned on a reaper thread: freeing an LP's large
    // vectors gives pages back to the system, which on the scheduling
    // thread would hold up the next launch.
    void release(int64_t ticket) {
        drain_incoming();
        if (ticket < 0 || ticket >= int64_t(jobs_.size()) || !jobs_[ticket]) throw std::runtime_error("Unknown ticket");
        if (!jobs_[ticket]->finished) throw std::runtime_error("Ticket has not finished");
        reaper_.give(std::move(jobs_[ticket]));
    }

    bool result(int64_t ticket, niryukti_result *out) {
        drain_incoming();
        if (ticket < 0 || ticket >= int64_t(jobs_.size())) throw std::runtime_error("Unknown ticket");
        if (!jobs_[ticket]) throw std::runtime_error("Ticket was released");
        Job &j = *jobs_[ticket];
        if (!j.finished) return false;
        const auto &r = j.result;
        const auto &a = r.accuracy;
        std::snprintf(out->status, sizeof out->status, "%s", r.status.c_str());
        std::snprintf(out->message, sizeof out->message, "%s", r.message.c_str());
        out->objective = a.finite ? j.original.sense * a.objective : NAN;
        out->primal_residual = a.primal;
        out->dual_residual = a.dual;
        out->relative_gap = a.gap;
        out->kkt_error = a.kkt;
        out->complementarity = a.complementarity;
        out->dual_bound = a.lower_bound;
        out->primal_absolute = a.primal_absolute;
        out->dual_absolute = a.dual_absolute;
        out->finite = a.finite;
        out->iterations = r.iterations;
        out->restarts = r.restarts;
        out->rejected_steps = r.rejected_steps;
        out->weight_updates = r.weight_updates;
        out->host_candidate_checks = r.host_candidate_checks;
        out->removed_rows = r.removed_rows;
        out->removed_columns = r.removed_columns;
        out->seconds = r.seconds;
        out->preprocess_seconds = r.preprocess_seconds;
        out->transfer_seconds = r.transfer_seconds;
        out->iteration_seconds = r.iteration_seconds;
        out->verification_seconds = r.verification_seconds;
        out->index_bits = 32;
        out->device_prepare = 0;
        out->operator_norm_estimate = r.operator_norm_estimate;
        out->polishing_iterations = 0;
        out->polishing_attempts = 0;
        out->initial_step = r.initial_step;
        out->initial_weight = r.initial_weight;
        out->first_restart_weight = r.first_restart_weight;
        out->first_restart_iteration = r.first_restart_iteration;
        out->fp32_iterations = r.fp32_iterations;
        if (out->x && r.x.size() == size_t(j.original.cols())) std::memcpy(out->x, r.x.data(), r.x.size() * 8);
        if (out->y && r.y.size() == size_t(j.original.rows())) std::memcpy(out->y, r.y.data(), r.y.size() * 8);
        return true;
    }

    void stats(niryukti_batch_stats *s) {
        drain_incoming();
        s->submitted = int64_t(jobs_.size());
        s->queued = int64_t(queue_.size()) + prep_outstanding_;
        s->active = 0;
        for (int v : slot_job_) s->active += v >= 0;
        s->finished = 0;
        for (const auto &j : jobs_) s->finished += !j || j->finished;
        s->arena_bytes = int64_t(arena_.size());
        {
            std::lock_guard<std::mutex> lock(arena_mu_);
            s->arena_used_bytes = int64_t(arena_.used());
        }
        s->launches = launches_;
        s->host_certifications = certifications_;
        s->host_farkas_checks = farkas_checks_;
        s->handoffs = handoffs_total_;
        s->handoffs_running = handoffs_outstanding_;
        s->device_seconds = device_seconds_;
        s->advance_seconds = advance_seconds_;
        s->check_seconds = check_seconds_;
        s->advance_bytes = advance_bytes_;
        s->attempts = attempts_;
        for (int c = 0; c <= kClasses; ++c) {
            s->class_seconds[c] = class_seconds_[c];
            s->class_bytes[c] = class_bytes_[c];
            s->class_launches[c] = class_launches_[c];
        }
        for (int c = 0; c < kClasses - 1; ++c) {
            const Shape sh = kShapes[tuner_.next(c)];
            s->class_threads[c] = sh.threads;
            s->class_rows[c] = sh.rows;
            s->class_tuned[c] = tuner_.decided(c) ? 1 : 0;
        }
        s->pad = 0;
        s->certify_wait_seconds = certify_wait_seconds_;
        s->settle_seconds = settle_seconds_;
        s->admit_seconds = admit_seconds_;
        for (int k = 0; k < kShapeCount; ++k) {
            s->shape_seconds[k] = shape_seconds_[k];
            s->shape_bytes[k] = shape_bytes_[k];
            s->shape_launches[k] = shape_launches_[k];
        }
    }

    // Waits for every submitted LP's prepare, then admits what the slots hold.
    int64_t admit_only() {
        while (prep_outstanding_ > 0) collect_prepared(true);
        int64_t before = 0, after = 0;
        for (int v : slot_job_) before += v >= 0;
        admit();
        for (int v : slot_job_) after += v >= 0;
        return after - before;
    }

  private:
    // Host part of src/solver.cpp:215-333 for one LP: prepare, start point,
    // start-point verification, unboundedness and trivial optimality, steps.
    // Runs on a prepare worker; returns true when the result is decided
    // without iterating (the scheduling thread then marks the job finished).
    bool setup(Job &j, const nyk::Options &o) {
        auto &r = j.result;
        r.method_selected = "pdhg";
        r.backend = "hip-batch";
        j.verifier = std::make_unique<nyk::Verifier>(j.original);
        r.status = "ITERATION_LIMIT";
        auto t0 = Clock::now();
        j.prep = nyk::prepare(j.original, o);
        r.preprocess_seconds = std::chrono::duration<double>(Clock::now() - t0).count();
        auto finish = [&](const char *status, std::string message) {
            r.status = status;
            r.message = std::move(message);
            r.seconds = std::chrono::duration<double>(Clock::now() - j.start).count();
            return true;
        };
        if (!j.prep.failure.empty()) {
            r.infeasibility_ray = j.prep.failure_ray;
            if (!r.infeasibility_ray.empty()) r.certificate_margin = j.verifier->infeasibility_bound(r.infeasibility_ray);
            return finish(j.prep.failure.c_str(), j.prep.reason);
        }
        const nyk::Model &m = j.prep.model;
        r.removed_columns = j.original.cols() - m.cols();
        r.removed_rows = j.original.rows() - m.rows();
        j.x.assign(m.cols(), 0);
        j.y.assign(m.rows(), 0);
        for (int64_t c = 0; c < m.cols(); ++c) {
            double initial = o.initial_x.empty() ? 0 : o.initial_x[j.prep.cols[c]];
            j.x[c] = std::clamp(initial / j.prep.column_scale[c], m.lb[c], m.ub[c]);
        }
        if (!o.initial_y.empty())
            for (int64_t i = 0; i < m.rows(); ++i)
                j.y[i] = o.initial_y[j.prep.rows[i]] / (j.prep.row_scale[i] * j.prep.objective_scale);
        r.x = j.prep.restore_x(j.x);
        r.y = j.prep.restore_y(j.y, j.original.rl.size());
        r.accuracy = j.verifier->evaluate(r.x, r.y);
        if (r.accuracy.finite && r.accuracy.primal_absolute == 0) {
            std::vector<uint8_t> used(m.cols(), 0);
            for (auto c : m.A.index) used[c] = 1;
            for (int64_t c = 0; c < m.cols(); ++c)
                if (!used[c] && ((m.c[c] < 0 && m.ub[c] == nyk::kInf) || (m.c[c] > 0 && m.lb[c] == -nyk::kInf)))
                    return finish("UNBOUNDED",
                                  "Feasible point and isolated improving recession direction: " + m.column_name(c));
        }
        if (r.accuracy.finite && r.accuracy.kkt <= o.tol) return finish("OPTIMAL", "");
        if (o.time_limit == 0) return finish("TIME_LIMIT", "");
        const double norm = nyk::step_norm(m, o, &j.result.operator_norm_estimate);
        j.step = norm > 0 ? .9 / norm : 1;
        j.weight = 1;
        if (o.adaptive) {
            long double c2 = 0, b2 = 0;
            for (double c : m.c) c2 += (long double)c * c;
            for (int64_t i = 0; i < m.rows(); ++i) {
                double b = std::clamp(0., m.rl[i], m.ru[i]);
                b2 += (long double)b * b;
            }
            if (c2 > 1e-24 && b2 > 1e-24) j.weight = std::clamp(std::sqrt(double(c2 / b2)), 1e-4, 1e4);
        }
        r.initial_step = j.step;
        r.initial_weight = j.weight;
        sort_by_length(j.prep, j.x, j.y);
        return false;
    }

    // Reorders the prepared LP's rows and columns by decreasing length (a
    // stable sort; the step and weight above are already taken), so that the
    // 64-row slices of the sliced copies hold rows of similar length: the
    // padding a slice carries to its longest row falls from about 1.5-1.8
    // times the nonzeros to a few percent on the batched workload, and with
    // it the loads of every step.  Each row keeps its entries in their order
    // (so every row's dot product is the same sum), the prepare maps are
    // composed with the reordering (restore_x / restore_y, the device's
    // restore and certification see the original LP), and the start point is
    // reordered with the LP.  The order depends only on the LP, so a result
    // is the same alone and in any mix.
    static void sort_by_length(nyk::Prepared &p, std::vector<double> &x, std::vector<double> &y) {
        nyk::Model &m = p.model;
        const int64_t rows = m.rows(), cols = m.cols();
        std::vector<int64_t> col_len(static_cast<size_t>(cols), 0);
        for (int64_t c : m.A.index) ++col_len[size_t(c)];
        std::vector<int64_t> rperm(static_cast<size_t>(rows)), cperm(static_cast<size_t>(cols));
        for (int64_t i = 0; i < rows; ++i) rperm[size_t(i)] = i;
        for (int64_t c = 0; c < cols; ++c) cperm[size_t(c)] = c;
        std::stable_sort(rperm.begin(), rperm.end(), [&](int64_t a, int64_t b) {
            return m.A.ptr[size_t(a) + 1] - m.A.ptr[size_t(a)] > m.A.ptr[size_t(b) + 1] - m.A.ptr[size_t(b)];
        });
        std::stable_sort(cperm.begin(), cperm.end(),
                         [&](int64_t a, int64_t b) { return col_len[size_t(a)] > col_len[size_t(b)]; });
        std::vector<int64_t> cnew(static_cast<size_t>(cols));
        for (int64_t c = 0; c < cols; ++c) cnew[size_t(cperm[size_t(c)])] = c;
        nyk::Csr a;
        a.rows = rows;
        a.cols = cols;
        a.ptr.assign(size_t(rows) + 1, 0);
        a.index.reserve(m.A.index.size());
        a.value.reserve(m.A.value.size());
        for (int64_t r = 0; r < rows; ++r) {
            const int64_t i = rperm[size_t(r)];
            for (int64_t k = m.A.ptr[size_t(i)]; k < m.A.ptr[size_t(i) + 1]; ++k) {
                a.index.push_back(cnew[size_t(m.A.index[size_t(k)])]);
                a.value.push_back(m.A.value[size_t(k)]);
            }
            a.ptr[size_t(r) + 1] = int64_t(a.index.size());
        }
        m.A = std::move(a);
        auto by_col = [&](auto &v) {
            auto old = v;
            for (int64_t c = 0; c < cols; ++c) v[size_t(c)] = old[size_t(cperm[size_t(c)])];
        };
        auto by_row = [&](auto &v) {
            auto old = v;
            for (int64_t r = 0; r < rows; ++r) v[size_t(r)] = old[size_t(rperm[size_t(r)])];
        };
        by_col(m.c);
        by_col(m.lb);
        by_col(m.ub);
        by_col(p.column_scale);
        by_col(p.cols);
        by_col(x);
        by_row(m.rl);
        by_row(m.ru);
        by_row(p.row_scale);
        by_row(p.rows);
        by_row(p.lower_rows);
        by_row(p.upper_rows);
        by_row(y);
        if (!m.names.empty()) by_col(m.names);
        if (!m.row_names.empty()) by_row(m.row_names);
    }

    // Algorithmic HBM bytes of one PDHG attempt on the scaled LP: both CSR
    // streams (8-byte values, 4-byte indices and offsets), each gathered
    // vector once, and eight FP64 vector streams per side (read x, xa, c, lb,
    // ub / write x, xa, dx; read y, ya, ax, rl, ru / write y, ya, ax).
    // (an entry is 8 value bytes and a 16-bit index in the sliced copies of
    // the LDS classes, a 32-bit index in the global class's CSR)
    // Algorithmic bytes of one attempt: each entry of A and A^T (value and
    // index: 8 + 2 in the sliced layouts, 4 + 2 with FP32 values, 8 + 4 in
    // CSR), the offsets, one read of y and dx, and 64 bytes of vector
    // streams per row and column.
    double attempt_bytes(const Job &j, bool fp32 = false) const {
        const double n = double(j.prep.model.cols()), m = double(j.prep.model.rows()), nnz = double(j.prep.model.A.nnz());
        const double entry = j.cls < kClasses - 1 ? (fp32 ? 6 : 10) : 12;
        return 2 * nnz * entry + (n + m + 2) * 4 + m * 8 + n * 8 + 64 * (n + m);
    }

    template <int NT, int NR>
    void launch_lds_t(int cap, int grid, hipStream_t cs, const int *slots, const BatchOpts &bo, bool fp32) {
        if (fp32)
            batch_advance_lds_kernel<NT, NR, float><<<grid, NT, lds_bytes(cap), cs>>>(d_desc_, d_state_, slots, bo);
        else
            batch_advance_lds_kernel<NT, NR, double><<<grid, NT, lds_bytes(cap), cs>>>(d_desc_, d_state_, slots, bo);
    }
    // Algorithmic bytes of one check period's check: the scaled SpMV that
    // rebases A x, then both candidates' original-space evaluations (restore,
    // one pass over the original A and one over its transpose for the two
    // candidates' products, and the row and column passes of the verifier).
    static double check_bytes(const Job &j) {
        const nyk::Model &s = j.prep.model, &o = j.original;
        const double n = double(s.cols()), m = double(s.rows()), nnz = double(s.A.nnz());
        const double n0 = double(o.cols()), m0 = double(o.rows()), nnz0 = double(o.A.nnz());
        return nnz * 12 + m * 12 + n * 8 + 24 * nnz0 + 2 * (68 * m0 + 100 * n0 + 28 * n + 32 * m);
    }

    // The tail group: one launch per scheduling round, one row per thread
    // over 1,024 threads (2 or 4 rows per thread spill registers at this size
    // and were measured slower).  Each LP runs at least `checks` periods and
    // then continues until `budget` ticks have passed.  The LPs come from the
    // device-side queue at d_queue_ (zeroed with the round's uploads), so the
    // grid needs no more workgroups than the device runs at once: one
    // 1,024-thread workgroup per CU.
    void launch_tail(int cap, int count, hipStream_t cs, const int *slots, const BatchOpts &bo, int checks,
                     long long budget) {
        const int grid = std::max(1, std::min(count, compute_units_));
        batch_tail_kernel<kTailThreads><<<grid, kTailThreads, lds_bytes(cap), cs>>>(
            d_desc_, d_state_, slots, count, d_queue_, bo, checks, checks * kTailMaxPeriodsPerCheck, budget);
    }
    // A crowded LDS class in the launch shape its tuner picked.
    void launch_lds(int cap, int shape, int grid, hipStream_t cs, const int *slots, const BatchOpts &bo, bool fp32) {
        switch (shape) {
        case 3: launch_lds_t<kWideThreads, 1>(cap, grid, cs, slots, bo, fp32); break;
        case 2: launch_lds_t<kThreads, 4>(cap, grid, cs, slots, bo, fp32); break;
        case 1: launch_lds_t<kThreads, 2>(cap, grid, cs, slots, bo, fp32); break;
        default: launch_lds_t<kThreads, 1>(cap, grid, cs, slots, bo, fp32);
        }
    }

    // Dynamic LDS of an LP with m + n <= cap (lds_layout): y and dx
    // (doubles), the two slice-base arrays (ceil(n/64) + 1 and ceil(m/64) + 1
    // ints) and the two row-length arrays (n and m bytes).
    static size_t lds_bytes(int cap) { return size_t(cap) * 9 + size_t(cap / 64 + 6) * 4 + 16; }

    int lds_class(const Job &j) const {
        const nyk::Model &m = j.prep.model;
        int64_t longest = 0;
        for (int64_t i = 0; i < m.rows(); ++i) longest = std::max(longest, m.A.ptr[i + 1] - m.A.ptr[i]);
        std::vector<int64_t> col(m.cols(), 0);
        for (auto c : m.A.index) longest = std::max(longest, ++col[c]);
        const int64_t need = m.rows() + m.cols();
        if (longest <= kShortRow)
            for (int c = 0; c < kClasses - 1; ++c)
                if (need <= class_cap_[c]) return c;
        return kClasses - 1;
    }

    // Admits prepared LPs (already resident in the arena) into free slots (FIFO).
    void admit() {
        collect_prepared(false);
        while (!queue_.empty() && !free_slots_.empty()) {
            Job &j = *jobs_[queue_.front()];
            queue_.pop_front();
            if (std::chrono::duration<double>(Clock::now() - j.start).count() >= opt_.time_limit) {
                release_arena(j.offset, j.bytes);
                j.resident = false;
                retire_host(j, "TIME_LIMIT");
                continue;
            }
            const int s = free_slots_.back();
            free_slots_.pop_back();
            j.slot = s;
            slot_job_[s] = int(j.ticket);
            slot_bytes_[s] = j.attempt_b;
            slot_bytes32_[s] = j.attempt_b32;
            slot_check_bytes_[s] = j.check_b;
            slot_class_[s] = j.cls;
            desc_[s] = j.dev;
            LpState st{};
            st.step = j.step;
            st.weight = j.weight;
            st.factor = 1;
            st.restart_kkt = j.result.accuracy.kkt;
            st.best_kkt = j.result.accuracy.kkt;
            st.iteration_limit = opt_.iteration_limit;
            st.status = kRunning;
            st.stall_kkt = INFINITY;
            st.fp64 = mixed_ && j.cls < kClasses - 1 ? 0 : 1;  // the global class steps with FP64 values
            state_[s] = st;
            dirty_ = true;
        }
    }

    // Lays out and fills the device image of one LP; `place(total bytes)`
    // returns the device address the image will be copied to, which the
    // descriptor's pointers refer to.
    template <class Place>
    size_t build_image(const Job &j, std::vector<char> &image, LpDev &d, Place &&place) {
        const nyk::Model &m = j.prep.model;
        const nyk::Model &m0 = j.original;
        const nyk::Csr &t0 = j.verifier->transpose();
        const int n = int(m.cols()), mr = int(m.rows()), n0 = int(m0.cols()), mr0 = int(m0.rows());
        const int64_t nnz = m.A.nnz(), nnz0 = m0.A.nnz();
        Layout L;
        size_t o_ap = L.add<int>(mr + 1), o_ai = L.add<int>(nnz), o_av = L.add<double>(nnz);
        size_t o_tp = L.add<int>(n + 1), o_ti = L.add<int>(nnz), o_tv = L.add<double>(nnz);
        size_t o_c = L.add<double>(n), o_lb = L.add<double>(n), o_ub = L.add<double>(n);
        size_t o_rl = L.add<double>(mr), o_ru = L.add<double>(mr);
        size_t o_op = L.add<int>(mr0 + 1), o_oi = L.add<int>(nnz0), o_ov = L.add<double>(nnz0);
        size_t o_otp = L.add<int>(n0 + 1), o_oti = L.add<int>(nnz0), o_otv = L.add<double>(nnz0);
        size_t o_c0 = L.add<double>(n0), o_lb0 = L.add<double>(n0), o_ub0 = L.add<double>(n0);
        size_t o_rl0 = L.add<double>(mr0), o_ru0 = L.add<double>(mr0);
        size_t o_cols = L.add<int>(n), o_lr = L.add<int>(mr), o_ur = L.add<int>(mr);
        size_t o_fixed = L.add<double>(n0), o_dc = L.add<double>(n), o_dr = L.add<double>(mr);
        size_t o_x[2], o_xa[2], o_y[2], o_ya[2], o_ax[2];
        for (int h = 0; h < 2; ++h) {
            o_x[h] = L.add<double>(n);
            o_xa[h] = L.add<double>(n);
            o_y[h] = L.add<double>(mr);
            o_ya[h] = L.add<double>(mr);
            o_ax[h] = L.add<double>(mr);
        }
        size_t o_dx = L.add<double>(n), o_epx = L.add<double>(n), o_epy = L.add<double>(mr);
        size_t o_ox[2], o_oy[2];
        for (int h = 0; h < 2; ++h) {
            o_ox[h] = L.add<double>(n0);
            o_oy[h] = L.add<double>(mr0);
        }
        size_t o_act = L.add<double>(mr0), o_aty = L.add<double>(n0), o_atyabs = L.add<double>(n0);
        size_t o_act1 = L.add<double>(mr0), o_aty1 = L.add<double>(n0), o_atyabs1 = L.add<double>(n0);
        size_t o_bx = L.add<double>(n0), o_by = L.add<double>(mr0);
        const auto ba = row_blocks(m.A.ptr), bt = row_blocks(t_ptr(m));
        size_t o_ab = L.add<int>(ba.size()), o_tb = L.add<int>(bt.size());
        const auto la = long_rows(m.A.ptr), lo = long_rows(m0.A.ptr), lot = long_rows(t0.ptr);
        size_t o_la = L.add<int>(la.size()), o_lo = L.add<int>(lo.size()), o_lot = L.add<int>(lot.size());
        // the sliced copies serve only the LDS classes (the global class walks the CSR)
        Sliced sa, st;
        nyk::Csr t = m.A.transposed();
        const bool fp32 = mixed_ && j.cls < kClasses - 1;
        if (j.cls < kClasses - 1) {
            sa = slice_rows(m.A.ptr, m.A.index, m.A.value, fp32);
            st = slice_rows(t.ptr, t.index, t.value, fp32);
        }
        size_t o_asb = L.add<int>(sa.base.size()), o_asi = L.add<unsigned short>(sa.index.size()),
               o_asv = L.add<double>(sa.value.size());
        size_t o_tsb = L.add<int>(st.base.size()), o_tsi = L.add<unsigned short>(st.index.size()),
               o_tsv = L.add<double>(st.value.size());
        const size_t o_asv32 = fp32 ? L.add<float>(sa.value32.size()) : 0,
                     o_tsv32 = fp32 ? L.add<float>(st.value32.size()) : 0;
        const bool lds = j.cls < kClasses - 1;
        const size_t o_lub = lds ? L.add<double>(2 * size_t(n)) : 0, o_rlu = lds ? L.add<double>(2 * size_t(mr)) : 0;
        char *const base = place(L.total);
        image.assign(L.total, 0);
        auto put_i = [&](size_t off, const std::vector<int64_t> &v) {
            int *p = reinterpret_cast<int *>(image.data() + off);
            for (size_t k = 0; k < v.size(); ++k) p[k] = int(v[k]);
        };
        auto put_d = [&](size_t off, const std::vector<double> &v) {
            if (!v.empty()) std::memcpy(image.data() + off, v.data(), v.size() * 8);
        };
        put_i(o_ap, m.A.ptr);
        put_i(o_ai, m.A.index);
        put_d(o_av, m.A.value);
        put_i(o_tp, t.ptr);
        put_i(o_ti, t.index);
        put_d(o_tv, t.value);
        put_d(o_c, m.c);
        put_d(o_lb, m.lb);
        put_d(o_ub, m.ub);
        put_d(o_rl, m.rl);
        put_d(o_ru, m.ru);
        put_i(o_op, m0.A.ptr);
        put_i(o_oi, m0.A.index);
        put_d(o_ov, m0.A.value);
        put_i(o_otp, t0.ptr);
        put_i(o_oti, t0.index);
        put_d(o_otv, t0.value);
        put_d(o_c0, m0.c);
        put_d(o_lb0, m0.lb);
        put_d(o_ub0, m0.ub);
        put_d(o_rl0, m0.rl);
        put_d(o_ru0, m0.ru);
        put_i(o_cols, j.prep.cols);
        put_i(o_lr, j.prep.lower_rows);
        put_i(o_ur, j.prep.upper_rows);
        put_d(o_fixed, j.prep.fixed);
        put_d(o_dc, j.prep.column_scale);
        put_d(o_dr, j.prep.row_scale);
        put_d(o_x[0], j.x);
        put_d(o_y[0], j.y);
        put_d(o_epx, j.x);
        put_d(o_epy, j.y);
        std::vector<double> ax0(mr, 0);
        m.A.multiply(j.x.data(), ax0.data());
        put_d(o_ax[0], ax0);
        put_d(o_bx, j.result.x);
        put_d(o_by, j.result.y);
        auto P = [&](size_t off) { return base + off; };
        d.n = n;
        d.m = mr;
        d.n0 = n0;
        d.m0 = mr0;
        d.ap = (const int *)P(o_ap);
        d.ai = (const int *)P(o_ai);
        d.av = (const double *)P(o_av);
        d.tp = (const int *)P(o_tp);
        d.ti = (const int *)P(o_ti);
        d.tv = (const double *)P(o_tv);
        d.c = (const double *)P(o_c);
        d.lb = (const double *)P(o_lb);
        d.ub = (const double *)P(o_ub);
        d.rl = (const double *)P(o_rl);
        d.ru = (const double *)P(o_ru);
        d.op = (const int *)P(o_op);
        d.oi = (const int *)P(o_oi);
        d.ov = (const double *)P(o_ov);
        d.otp = (const int *)P(o_otp);
        d.oti = (const int *)P(o_oti);
        d.otv = (const double *)P(o_otv);
        d.c0 = (const double *)P(o_c0);
        d.lb0 = (const double *)P(o_lb0);
        d.ub0 = (const double *)P(o_ub0);
        d.rl0 = (const double *)P(o_rl0);
        d.ru0 = (const double *)P(o_ru0);
        d.cols = (const int *)P(o_cols);
        d.lower_rows = (const int *)P(o_lr);
        d.upper_rows = (const int *)P(o_ur);
        d.fixed = (const double *)P(o_fixed);
        d.dc = (const double *)P(o_dc);
        d.dr = (const double *)P(o_dr);
        d.sobj = j.prep.objective_scale;
        d.offset0 = m0.offset;
        double cscale = 1;
        for (double v : m0.c) cscale = std::max(cscale, 1 + std::abs(v));
        d.cscale = cscale;
        for (int h = 0; h < 2; ++h) {
            d.x[h] = (double *)P(o_x[h]);
            d.xa[h] = (double *)P(o_xa[h]);
            d.y[h] = (double *)P(o_y[h]);
            d.ya[h] = (double *)P(o_ya[h]);
            d.ax[h] = (double *)P(o_ax[h]);
            d.ox[h] = (double *)P(o_ox[h]);
            d.oy[h] = (double *)P(o_oy[h]);
        }
        d.dx = (double *)P(o_dx);
        d.epx = (double *)P(o_epx);
        d.epy = (double *)P(o_epy);
        d.act = (double *)P(o_act);
        d.aty = (double *)P(o_aty);
        d.atyabs = (double *)P(o_atyabs);
        d.act1 = (double *)P(o_act1);
        d.aty1 = (double *)P(o_aty1);
        d.atyabs1 = (double *)P(o_atyabs1);
        d.bx = (double *)P(o_bx);
        d.by = (double *)P(o_by);
        put_i(o_ab, ba);
        put_i(o_tb, bt);
        d.ablk = (const int *)P(o_ab);
        d.tblk = (const int *)P(o_tb);
        put_i(o_la, la);
        put_i(o_lo, lo);
        put_i(o_lot, lot);
        d.along = (const int *)P(o_la);
        d.olong = (const int *)P(o_lo);
        d.otlong = (const int *)P(o_lot);
        d.nalong = int(la.size());
        d.nolong = int(lo.size());
        d.notlong = int(lot.size());
        auto put_u16 = [&](size_t off, const std::vector<int64_t> &v) {
            unsigned short *p = reinterpret_cast<unsigned short *>(image.data() + off);
            for (size_t k = 0; k < v.size(); ++k) p[k] = (unsigned short)v[k];
        };
        put_i(o_asb, sa.base);
        put_u16(o_asi, sa.index);
        put_d(o_asv, sa.value);
        put_i(o_tsb, st.base);
        put_u16(o_tsi, st.index);
        put_d(o_tsv, st.value);
        d.asv32 = d.tsv32 = nullptr;
        if (fp32) {
            if (!sa.value32.empty()) std::memcpy(image.data() + o_asv32, sa.value32.data(), sa.value32.size() * 4);
            if (!st.value32.empty()) std::memcpy(image.data() + o_tsv32, st.value32.data(), st.value32.size() * 4);
            d.asv32 = (const float *)P(o_asv32);
            d.tsv32 = (const float *)P(o_tsv32);
        }
        d.lub = d.rlu = nullptr;
        if (lds) {
            double *lub = reinterpret_cast<double *>(image.data() + o_lub);
            double *rlu = reinterpret_cast<double *>(image.data() + o_rlu);
            for (int jj = 0; jj < n; ++jj) {
                lub[2 * jj] = m.lb[size_t(jj)];
                lub[2 * jj + 1] = m.ub[size_t(jj)];
            }
            for (int i = 0; i < mr; ++i) {
                rlu[2 * i] = m.rl[size_t(i)];
                rlu[2 * i + 1] = m.ru[size_t(i)];
            }
            d.lub = (const double *)P(o_lub);
            d.rlu = (const double *)P(o_rlu);
        }
        d.asb = (const int *)P(o_asb);
        d.asi = (const unsigned short *)P(o_asi);
        d.asv = (const double *)P(o_asv);
        d.tsb = (const int *)P(o_tsb);
        d.tsi = (const unsigned short *)P(o_tsi);
        d.tsv = (const double *)P(o_tsv);
        d.nab = int(ba.size()) - 1;
        d.ntb = int(bt.size()) - 1;
        return L.total;
    }

    // Queues a device-to-host copy; flush_downloads() performs every queued
    // copy with one gather kernel and one transfer (thousands of small copies
    // per scheduling round otherwise cost a runtime call each).
    void download(double *host, const double *dev, size_t count) {
        if (count) pending_.push_back(Pending{host, dev, count});
    }
    void flush_downloads() {
        if (pending_.empty()) {
            HIP_OK(hipStreamSynchronize(stream_));
            return;
        }
        std::vector<CopyPiece> pieces;
        pieces.reserve(pending_.size());
        long long total = 0;
        for (const Pending &p : pending_) {
            pieces.push_back(CopyPiece{p.dev, total, (long long)p.count});
            total += (long long)p.count;
        }
        if (total > stage_capacity_) {
            if (d_stage_) HIP_OK(hipFree(d_stage_));
            if (h_stage_) HIP_OK(hipHostFree(h_stage_));
            d_stage_ = nullptr;
            h_stage_ = nullptr;
            stage_capacity_ = std::max<long long>(total, 2 * stage_capacity_);
            HIP_OK(hipMalloc(&d_stage_, size_t(stage_capacity_) * 8));
            HIP_OK(hipHostMalloc(&h_stage_, size_t(stage_capacity_) * 8, hipHostMallocDefault));
        }
        if ((long long)pieces.size() > piece_capacity_) {
            if (d_pieces_) HIP_OK(hipFree(d_pieces_));
            d_pieces_ = nullptr;
            piece_capacity_ = std::max<long long>((long long)pieces.size(), 2 * piece_capacity_);
            HIP_OK(hipMalloc(&d_pieces_, sizeof(CopyPiece) * size_t(piece_capacity_)));
        }
        HIP_OK(hipMemcpyAsync(d_pieces_, pieces.data(), sizeof(CopyPiece) * pieces.size(), hipMemcpyHostToDevice,
                              stream_));
        gather_pieces_kernel<<<int(pieces.size()), kThreads, 0, stream_>>>(d_pieces_, d_stage_);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(h_stage_, d_stage_, size_t(total) * 8, hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
        for (size_t k = 0; k < pending_.size(); ++k)
            std::memcpy(pending_[k].host, h_stage_ + pieces[k].offset, pending_[k].count * 8);
        pending_.clear();
    }

    // Handles every LP whose workgroup stopped: certification, Farkas checks,
    // limits and errors (src/solver.cpp:583-633, 727-731).
    // LPs that stopped in one round.  The host certifies them on a thread of
    // its own while the next round runs on the device (their slots stay
    // parked, so the device skips them), and the outcome is applied after
    // that round: a finished LP is retired one round later, an LP whose
    // certification fails resumes one round later.  Neither changes any
    // result, only when it is reported.
    struct Settling {
        std::vector<int> slots;
        std::vector<Job *> jobs;
        std::vector<LpState> states;  // the parked states, as downloaded
        std::vector<std::vector<double>> cx, cyo, cy, cya, epy;
        std::vector<int> resume;  // 0 finished, 1 resume, 2 hand off
        std::vector<double> cand_kkt;
        int64_t certifications = 0, farkas_checks = 0;
        nyk::Beside task;  // the certification, when it runs beside the next round
    };

    // Collects the LPs that stopped in the round just finished (including
    // those finish_round stopped at their time limit) and downloads what
    // their certification needs.
    std::unique_ptr<Settling> collect(const std::vector<int> &active) {
        std::vector<int> stopped;
        for (int s : active)
            if (state_[s].status != kRunning) stopped.push_back(s);
        if (stopped.empty()) return nullptr;
        auto p = std::make_unique<Settling>();
        const size_t ns = stopped.size();
        p->slots = stopped;
        p->jobs.resize(ns);
        p->states.resize(ns);
        p->cx.resize(ns);
        p->cyo.resize(ns);
        p->cy.resize(ns);
        p->cya.resize(ns);
        p->epy.resize(ns);
        p->resume.assign(ns, 0);
        p->cand_kkt.assign(ns, 0.0);
        for (size_t k = 0; k < ns; ++k) {
            const int s = stopped[k];
            const LpDev &d = desc_[s];
            const LpState &st = state_[s];
            Job &j = *jobs_[slot_job_[s]];
            p->jobs[k] = &j;
            p->states[k] = st;
            if (st.status == kHandoff) {
                // the current (scaled) iterate; restored on the host below
                p->cx[k].resize(d.n);
                p->cyo[k].resize(d.m);
                download(p->cx[k].data(), d.x[st.cur], d.n);
                download(p->cyo[k].data(), d.y[st.cur], d.m);
            } else if (st.status == kParkOptimal) {
                // the parked candidate is still in its evaluation buffers
                p->cx[k].resize(d.n0);
                p->cyo[k].resize(d.m0);
                download(p->cx[k].data(), d.ox[st.avg], d.n0);
                download(p->cyo[k].data(), d.oy[st.avg], d.m0);
            } else {
                j.result.x.resize(d.n0);
                j.result.y.resize(d.m0);
                download(j.result.x.data(), d.bx, d.n0);
                download(j.result.y.data(), d.by, d.m0);
            }
            if (st.status == kParkFarkas) {
                const int c = st.cur;
                p->cy[k].resize(d.m);
                p->cya[k].resize(d.m);
                p->epy[k].resize(d.m);
                download(p->cy[k].data(), d.y[c], d.m);
                download(p->cya[k].data(), d.ya[c], d.m);
                download(p->epy[k].data(), d.epy, d.m);
            }
        }
        flush_downloads();
        return p;
    }

    // Host certification of a round's stopped LPs; touches only their jobs
    // and the Settling record (never state_, which the next round's
    // download rewrites meanwhile).
    void certify(Settling &p) {
        std::mutex mu;
        nyk::parallel_tasks(int64_t(p.slots.size()), [&](int64_t k) {
            Job &j = *p.jobs[k];
            const LpState &st = p.states[k];
            auto &r = j.result;
            auto t0 = Clock::now();
            copy_counters(j, st);
            if (st.status == kHandoff) {
                j.handoff_x = j.prep.restore_x(p.cx[k]);
                j.handoff_y = j.prep.restore_y(p.cyo[k], j.original.rl.size());
                p.resume[k] = 2;
            } else if (st.status == kParkOptimal) {
                // src/solver.cpp:588-596: OPTIMAL only after the safe re-evaluation
                auto safe = j.verifier->evaluate(p.cx[k], p.cyo[k]);
                {
                    std::lock_guard<std::mutex> lock(mu);
                    ++p.certifications;
                }
                if (safe.finite && safe.kkt <= opt_.tol) {
                    r.x = std::move(p.cx[k]);
                    r.y = std::move(p.cyo[k]);
                    r.accuracy = safe;
                    r.status = "OPTIMAL";
                } else {
                    p.cand_kkt[k] = safe.kkt;  // the reference continues with the safe accuracy
                    p.resume[k] = 1;
                }
            } else if (st.status == kParkFarkas) {
                {
                    std::lock_guard<std::mutex> lock(mu);
                    ++p.farkas_checks;
                }
                const bool avg = st.avg != 0;
                auto ysel = avg ? p.cya[k] : p.cy[k], yother = avg ? p.cy[k] : p.cya[k];
                auto disp = ysel;
                for (size_t i = 0; i < disp.size(); ++i) disp[i] -= p.epy[k][i];
                std::vector<std::vector<double>> rays{j.prep.restore_y(ysel, j.original.rl.size()),
                                                      j.prep.restore_y(yother, j.original.rl.size()),
                                                      j.prep.restore_y(disp, j.original.rl.size())};
                bool infeasible = false;
                for (auto &ray : rays) {
                    double norm_y = 0;
                    bool finite = true;
                    for (double v : ray) {
                        finite = finite && std::isfinite(v);
                        norm_y = std::max(norm_y, std::abs(v));
                    }
                    if (!finite || !(norm_y > 0)) continue;
                    for (double &v : ray) v /= norm_y;
                    double margin = j.verifier->infeasibility_bound(ray);
                    if (margin > 0 && std::isfinite(margin)) {
                        r.status = "INFEASIBLE";
                        r.message = "Independently verified box/row Farkas certificate";
                        r.infeasibility_ray = std::move(ray);
                        r.certificate_margin = margin;
                        infeasible = true;
                        break;
                    }
                }
                if (!infeasible) {
                    p.cand_kkt[k] = st.cand_kkt;
                    p.resume[k] = 1;
                }
            } else if (st.status == kNumerical) {
                r.status = "NUMERICAL_ERROR";
                r.message = "Nonfinite iterate or diagnostic";
            } else if (st.status == kExhausted) {
                r.status = "NUMERICAL_ERROR";
                r.message = "PDHG backtracking exhausted";
            } else if (r.status != "TIME_LIMIT") {
                r.status = "ITERATION_LIMIT";
            }
            r.verification_seconds += std::chrono::duration<double>(Clock::now() - t0).count();
            if (!p.resume[k]) {
                // final original-space verification (src/solver.cpp:727-731)
                if (r.status != "OPTIMAL") r.accuracy = j.verifier->evaluate(r.x, r.y);
                if (r.status == "OPTIMAL" && (!r.accuracy.finite || r.accuracy.kkt > opt_.tol)) r.status = "NUMERICAL_ERROR";
            }
        });
    }

    // Applies a finished certification: retire, resume (from the next
    // round on) or hand off each LP.
    void apply(Settling &p) {
        certifications_ += p.certifications;
        farkas_checks_ += p.farkas_checks;
        for (size_t k = 0; k < p.slots.size(); ++k) {
            Job &j = *p.jobs[k];
            if (p.resume[k] == 1) {
                LpState &st = state_[p.slots[k]];
                st.cand_kkt = p.cand_kkt[k];
                st.phase = 1;
                st.status = kRunning;
            } else if (p.resume[k] == 2) {
                hand_off(j);
            } else {
                retire(j);
            }
        }
    }

    // Waits for the previous round's certification and applies it.
    void finish_settling() {
        if (!settling_) return;
        std::unique_ptr<Settling> p = std::move(settling_);
        const auto t0 = Clock::now();
        p->task.wait();  // rethrows a certification failure
        certify_wait_seconds_ += std::chrono::duration<double>(Clock::now() - t0).count();
        apply(*p);
    }

    void settle(const std::vector<int> &active) {
        finish_settling();
        const auto t0 = Clock::now();
        std::unique_ptr<Settling> p = collect(active);
        settle_seconds_ += std::chrono::duration<double>(Clock::now() - t0).count();
        if (!p) return;
        if (p->slots.size() <= size_t(kInlineCertify)) {
            // a few LPs: certify now (a millisecond or so), so their results
            // are not held back a round (what a served request would wait)
            certify(*p);
            apply(*p);
            return;
        }
        Settling *raw = p.get();
        p->task.start([this, raw] { certify(*raw); });
        settling_ = std::move(p);
    }

    // Frees the LP's slot and arena space and queues it for a grid worker.
    void hand_off(Job &j) {
        release_arena(j.offset, j.bytes);
        j.resident = false;
        slot_job_[j.slot] = -1;
        free_slots_.push_back(j.slot);
        state_[j.slot] = LpState{};
        state_[j.slot].status = kIterationLimit;
        j.slot = -1;
        start_workers();
        {
            std::lock_guard<std::mutex> lock(handoff_mu_);
            handoff_queue_.push_back(&j);
        }
        ++handoffs_outstanding_;
        ++handoffs_total_;
        handoff_cv_.notify_one();
    }

    void start_workers() {
        if (!workers_.empty()) return;
        int device = 0;
        HIP_OK(hipGetDevice(&device));
        for (int w = 0; w < handoff_workers_; ++w)
            workers_.emplace_back([this, device] {
                (void)hipSetDevice(device);
                for (;;) {
                    Job *j = nullptr;  // jobs_ may grow meanwhile: work on the Job itself
                    {
                        std::unique_lock<std::mutex> lock(handoff_mu_);
                        handoff_cv_.wait(lock, [&] { return stopping_ || !handoff_queue_.empty(); });
                        if (handoff_queue_.empty()) return;
                        j = handoff_queue_.front();
                        handoff_queue_.pop_front();
                    }
                    run_handed_off(*j);
                    {
                        std::lock_guard<std::mutex> lock(handoff_mu_);
                        handoff_done_.push_back(j->ticket);
                    }
                    done_cv_.notify_all();
                }
            });
    }

    // Continues an LP on the grid-wide engine with the iteration and time
    // budget it has left; counters add up over both phases.
    void run_handed_off(Job &j) {
        auto &r = j.result;
        nyk::Options o = opt_;
        o.initial_x = std::move(j.handoff_x);
        o.initial_y = std::move(j.handoff_y);
        o.iteration_limit = std::max<int64_t>(0, opt_.iteration_limit - r.iterations);
        o.time_limit = std::max(0.0, opt_.time_limit - std::chrono::duration<double>(Clock::now() - j.start).count());
        o.device = "gpu";
        try {
            nyk::Result g = nyk::solve(j.original, o);
            const int64_t it = r.iterations, rs = r.restarts, rj = r.rejected_steps, wu = r.weight_updates;
            const int64_t hc = r.host_candidate_checks;
            const double pre = r.preprocess_seconds, ver = r.verification_seconds;
            r.status = g.status;
            r.message = g.message;
            r.x = std::move(g.x);
            r.y = std::move(g.y);
            r.accuracy = g.accuracy;
            r.infeasibility_ray = std::move(g.infeasibility_ray);
            r.certificate_margin = g.certificate_margin;
            r.iterations = it + g.iterations;
            r.restarts = rs + g.restarts;
            r.rejected_steps = rj + g.rejected_steps;
            r.weight_updates = wu + g.weight_updates;
            r.host_candidate_checks = hc + g.host_candidate_checks;
            r.device_candidate_checks = g.device_candidate_checks;
            r.preprocess_seconds = pre + g.preprocess_seconds;
            r.verification_seconds = ver + g.verification_seconds;
            r.iteration_seconds += g.iteration_seconds;
            r.backend = "hip-batch+hip";
        } catch (const std::exception &e) {
            r.status = "NUMERICAL_ERROR";
            r.message = std::string("Grid-engine continuation failed: ") + e.what();
        }
    }

    // Moves finished handoffs to the finished list (main thread only); waits
    // briefly for one when nothing else is running.
    void drain_handoffs(bool idle) {
        if (!handoffs_outstanding_) return;
        std::deque<int64_t> done;
        {
            std::unique_lock<std::mutex> lock(handoff_mu_);
            if (idle && handoff_done_.empty() && queue_.empty())
                done_cv_.wait_for(lock, std::chrono::milliseconds(2), [&] { return !handoff_done_.empty(); });
            done.swap(handoff_done_);
        }
        for (int64_t t : done) {
            Job &j = *jobs_[t];
            j.finished = true;
            done_.push_back(t);
            j.result.seconds = std::chrono::duration<double>(Clock::now() - j.start).count();
            --handoffs_outstanding_;
        }
    }

    void copy_counters(Job &j, const LpState &st) {
        auto &r = j.result;
        r.iterations = st.iterations;
        r.restarts = st.restarts;
        r.rejected_steps = st.rejected;
        r.weight_updates = st.weight_updates;
        r.host_candidate_checks = st.checks;
        r.first_restart_iteration = st.first_restart;
        r.first_restart_weight = st.first_weight;
        r.fp32_iterations = st.fp32_iterations;
    }

    void retire(Job &j) {
        release_arena(j.offset, j.bytes);
        j.resident = false;
        slot_job_[j.slot] = -1;
        free_slots_.push_back(j.slot);
        state_[j.slot] = LpState{};
        state_[j.slot].status = kIterationLimit;
        j.slot = -1;
        j.finished = true;
        done_.push_back(j.ticket);
        j.result.seconds = std::chrono::duration<double>(Clock::now() - j.start).count();
    }
    void retire_host(Job &j, const char *status) {
        j.result.status = status;
        j.finished = true;
        done_.push_back(j.ticket);
        j.result.seconds = std::chrono::duration<double>(Clock::now() - j.start).count();
    }

    nyk::Options opt_;
    int max_slots_ = 0;
    Arena arena_;
    hipStream_t stream_ = nullptr;
    LpDev *d_desc_ = nullptr;
    LpState *d_state_ = nullptr;
    int *d_slots_ = nullptr;
    unsigned *d_queue_ = nullptr;  // the tail launch's work-queue counter
    const unsigned zero_ = 0;
    std::vector<LpDev> desc_;
    std::vector<LpState> state_;
    std::vector<int> slot_job_, free_slots_;
    std::vector<double> slot_bytes_, slot_bytes32_, slot_check_bytes_;
    bool mixed_ = false;  // mixed precision: FP32 matrix values, then FP64
    std::unique_ptr<Settling> settling_;  // the certification running beside the next round
    long long tail_budget_ticks_ = 0;
    int compute_units_ = 0;
    hipStream_t class_stream_[kClasses + 1] = {};  // per class, then the tail group
    double *d_pow3_ = nullptr, *d_pow6_ = nullptr;
    struct Pending {
        double *host;
        const double *dev;
        size_t count;
    };
    std::vector<Pending> pending_;
    double *d_stage_ = nullptr, *h_stage_ = nullptr;
    CopyPiece *d_pieces_ = nullptr;
    long long stage_capacity_ = 0, piece_capacity_ = 0;
    long long pow_count_ = 0;
    ShapeTuner tuner_;
    int class_cap_[kClasses - 1] = {};
    std::vector<hipEvent_t> class_events_[kClasses + 1];
    hipEvent_t uploaded_ = nullptr;
    std::unique_ptr<Round> round_;  // launched, not yet finished
    std::vector<LpState> readback_;  // the device's state array after a round
    std::vector<LpState> upload_;    // the state array a launch uploads
    std::vector<int> slot_class_, slot_group_;
    double advance_seconds_ = 0, check_seconds_ = 0, advance_bytes_ = 0;
    double class_seconds_[kClasses + 1] = {}, class_bytes_[kClasses + 1] = {};  // per class, then the tail group
    int64_t class_launches_[kClasses + 1] = {};
    double shape_seconds_[kShapeCount] = {}, shape_bytes_[kShapeCount] = {};
    int64_t shape_launches_[kShapeCount] = {};
    int64_t attempts_ = 0;
    int device_ = 0;
    std::vector<std::unique_ptr<Job>> jobs_;  // scheduling thread only, indexed by ticket
    std::mutex submit_mu_;                           // the prepare workers' start
    nyk::Inbox<std::unique_ptr<Job>> incoming_;      // submitted jobs not yet in jobs_
    nyk::Reaper<std::unique_ptr<Job>> reaper_;       // released jobs, freed on its thread
    std::deque<int64_t> queue_, done_;
    // prepare workers: jobs to prepare, prepared jobs, and how many of the
    // submitted jobs the scheduling thread has not yet collected
    std::vector<std::thread> preparers_;
    std::mutex prep_mu_;
    mutable std::mutex arena_mu_;
    std::condition_variable prep_cv_, ready_cv_, arena_cv_;
    std::deque<Job *> prep_todo_;
    std::vector<Job *> prep_ready_;
    std::atomic<bool> prep_stop_{false};
    std::atomic<int64_t> prep_outstanding_{0};
    // handoff to the grid-wide engine
    int64_t handoff_ = 0;
    int handoff_workers_ = 4;
    int64_t handoffs_outstanding_ = 0, handoffs_total_ = 0;
    std::vector<std::thread> workers_;
    std::mutex handoff_mu_;
    std::condition_variable handoff_cv_, done_cv_;
    std::deque<Job *> handoff_queue_;
    std::deque<int64_t> handoff_done_;
    bool stopping_ = false;
    bool dirty_ = true;
    int64_t launches_ = 0, certifications_ = 0, farkas_checks_ = 0;
    double device_seconds_ = 0, certify_wait_seconds_ = 0, settle_seconds_ = 0, admit_seconds_ = 0;
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
    }
}

nyk::Options options_from(const niryukti_options *in) {
    nyk::Options o;
    if (!in) return o;
    auto str = [](const char *s, const char *d) { return std::string(s ? s : d); };
    o.device = str(in->device, "gpu");
    o.method = str(in->method, "pdhg");
    o.scaling = str(in->scaling, "ruiz");
    o.primal_weight = str(in->primal_weight, "displacement");
    o.matrix_precision = str(in->matrix_precision, "fp64");
    o.tol = in->tol;
    o.time_limit = in->time_limit;
    o.iteration_limit = in->iteration_limit;
    o.check_every = in->check_every;
    o.scaling_passes = in->scaling_passes;
    o.threads = in->threads;
    o.presolve = in->presolve != 0;
    o.restart = in->restart != 0;
    o.adaptive = in->adaptive != 0;
    o.power_iterations = in->power_iterations;
    o.polishing = in->polishing != 0;
    return o;
}

}  // namespace

struct niryukti_batch {
    BatchEngine engine;
    niryukti_batch(const nyk::Options &o, int64_t arena, int64_t slots) : engine(o, arena, slots) {}
};

extern "C" {

int niryukti_batch_create(const niryukti_options *opts, int64_t arena_bytes, int64_t max_slots, niryukti_batch **out) {
    return guarded([&] {
        nyk::Options o = options_from(opts);
        nyk::validate_options(o);
        if (o.primal_weight != "displacement" || (o.matrix_precision != "fp64" && o.matrix_precision != "mixed"))
            throw std::runtime_error("The batched engine supports the displacement weight and FP64 or mixed matrices");
        if (o.polishing)
            throw std::runtime_error("Feasibility polishing runs in the single-LP engine, not in the batched engine");
        *out = new niryukti_batch(o, arena_bytes, max_slots);
    });
}

void niryukti_batch_destroy(niryukti_batch *b) { delete b; }

int niryukti_batch_submit(niryukti_batch *b, const niryukti_lp *lp, const double *x0, const double *y0, int64_t *ticket) {
    return guarded([&] {
        if (!b || !lp) throw std::runtime_error("Null argument");
        int64_t t = b->engine.submit(*lp, x0, y0);
        if (ticket) *ticket = t;
    });
}

int niryukti_batch_step(niryukti_batch *b, int checks_per_launch, int64_t *running) {
    return guarded([&] {
        int64_t r = b->engine.step(std::max(1, checks_per_launch));
        if (running) *running = r;
    });
}

int niryukti_batch_admit(niryukti_batch *b, int64_t *admitted) {
    return guarded([&] {
        int64_t a = b->engine.admit_only();
        if (admitted) *admitted = a;
    });
}

int niryukti_batch_run(niryukti_batch *b, int checks_per_launch) {
    return guarded([&] {
        while (b->engine.step(std::max(1, checks_per_launch)) > 0) {
        }
    });
}

int niryukti_batch_result(niryukti_batch *b, int64_t ticket, niryukti_result *r) {
    int done = 0;
    int rc = guarded([&] { done = b->engine.result(ticket, r) ? 1 : 0; });
    return rc < 0 ? rc : done;
}

int niryukti_batch_set_handoff(niryukti_batch *b, int64_t iterations, int workers) {
    return guarded([&] {
        if (!b) throw std::runtime_error("Null batch");
        b->engine.set_handoff(iterations, workers);
    });
}

int niryukti_batch_finished(niryukti_batch *b, int64_t *tickets, int64_t capacity, int64_t *count) {
    return guarded([&] {
        if (!b || !count || (capacity > 0 && !tickets)) throw std::runtime_error("Null argument");
        *count = b->engine.finished(tickets, capacity);
    });
}

int niryukti_batch_collect(niryukti_batch *b, int64_t *tickets, niryukti_result *results, int64_t capacity,
                           int release, int64_t *count) {
    return guarded([&] {
        if (!b || !count || (capacity > 0 && (!tickets || !results))) throw std::runtime_error("Null argument");
        *count = b->engine.collect(tickets, results, capacity, release != 0);
    });
}

int niryukti_batch_release(niryukti_batch *b, int64_t ticket) {
    return guarded([&] {
        if (!b) throw std::runtime_error("Null batch");
        b->engine.release(ticket);
    });
}

int niryukti_batch_stats_get(niryukti_batch *b, niryukti_batch_stats *s) {
    return guarded([&] { b->engine.stats(s); });
}

const char *niryukti_batch_last_error(void) { return g_error.c_str(); }
}
