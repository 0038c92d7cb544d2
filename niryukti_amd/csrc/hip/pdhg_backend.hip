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

struct PrimalArgs {
    const double *c, *lb, *ub;
    double *x[2], *xa[2];
    double *dx;
    const double *y[2];
    const double *anchor_x;
    double *partial;
    State *st;
    Flags f;
};

struct DualArgs {
    const double *rl, *ru;
    double *y[2], *ya[2], *ax[2];
    const double *dx;
    const double *anchor_y, *anchor_ax;
    double *partial;  // 2 per block: |dy|^2, dy.delta
    const double *primal_partial;
    int primal_blocks;
    State *st;
    Flags f;
};

// ---------------------------------------------------------------- device --

template <int G>
__device__ __forceinline__ double group_sum(double v) {
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

template <int G, class O, class V>
__global__ __launch_bounds__(kBlock) void primal_kernel(DevCsr<O, V> at, PrimalArgs a) {
    const State *st = a.st;
    if (st->stop != kRun) return;
    const int cur = st->cur;
    const double *y = a.y[cur];
    const double *x = a.x[cur];
    double *xn = a.x[cur ^ 1];
    const double *xa = a.xa[cur];
    double *xan = a.xa[cur ^ 1];
    const double f = st->factor;
    const double tau = st->tau0 * f;
    const double w = f / (st->average_mass + f);
    const double lambda = 1 / (double(st->epoch_steps) + 2);
    const double rho = a.f.reflection;
    constexpr int rows_per_block = kBlock / G;
    const int lane = threadIdx.x % G;
    double local = 0;
    for (long long base = (long long)blockIdx.x * rows_per_block; base < at.rows;
         base += (long long)gridDim.x * rows_per_block) {
        const long long j = base + threadIdx.x / G;
        const bool valid = j < at.rows;
        double acc = 0;
        if (valid) {
            const O end = at.ptr[j + 1];
            for (O k = at.ptr[j] + lane; k < end; k += G) acc += double(at.val[k]) * y[at.idx[k]];
        }
        acc = group_sum<G>(acc);
        if (valid && lane == 0) {
            const double g = a.c[j] + acc;
            const double xj = x[j];
            const double t = clampd(xj - tau * g, a.lb[j], a.ub[j]);
            const double d = t - xj;
            local += d * d;
            a.dx[j] = d;
            if (a.f.halpern) {
                xan[j] = t;
                xn[j] = lambda * a.anchor_x[j] + (1 - lambda) * ((1 + rho) * t - rho * xj);
            } else {
                xn[j] = t;
                const double v = xa[j];
                xan[j] = v + (t - v) * w;
            }
        }
    }
    __shared__ double lds[4];
    const double s = block_sum(local, lds);
    if (threadIdx.x == 0) store_shared(&a.partial[blockIdx.x], s);
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

template <int G, class O, class V>
__global__ __launch_bounds__(kBlock) void dual_kernel(DevCsr<O, V> A, DualArgs a) {
    State *st = a.st;
    if (st->stop != kRun) return;
    const int cur = st->cur;
    const double *y = a.y[cur];
    double *yn = a.y[cur ^ 1];
    const double *ya = a.ya[cur];
    double *yan = a.ya[cur ^ 1];
    const double *ax = a.ax[cur];
    double *axn = a.ax[cur ^ 1];
    const double f = st->factor;
    const double sigma = st->sigma0 * f;
    const double w = f / (st->average_mass + f);
    const double lambda = 1 / (double(st->epoch_steps) + 2);
    const double rho = a.f.reflection;
    constexpr int rows_per_block = kBlock / G;
    const int lane = threadIdx.x % G;
    double sq = 0, cross = 0;
    for (long long base = (long long)blockIdx.x * rows_per_block; base < A.rows;
         base += (long long)gridDim.x * rows_per_block) {
        const long long i = base + threadIdx.x / G;
        const bool valid = i < A.rows;
        double acc = 0;
        if (valid) {
            const O end = A.ptr[i + 1];
            for (O k = A.ptr[i] + lane; k < end; k += G) acc += double(A.val[k]) * a.dx[A.idx[k]];
        }
        acc = group_sum<G>(acc);
        if (valid && lane == 0) {
            const double delta = acc;
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
                yn[i] = lambda * a.anchor_y[i] + (1 - lambda) * ((1 + rho) * t - rho * yi);
                axn[i] = lambda * a.anchor_ax[i] + (1 - lambda) * ((1 + rho) * next_ax - rho * axi);
            } else {
                yn[i] = t;
                const double v = ya[i];
                yan[i] = v + (t - v) * w;
                axn[i] = next_ax;
            }
        }
    }
    __shared__ double lds[4];
    __shared__ int last;
    const double s1 = block_sum(sq, lds);
    const double s2 = block_sum(cross, lds);
    if (threadIdx.x == 0) {
        store_shared(&a.partial[2 * blockIdx.x], s1);
        store_shared(&a.partial[2 * blockIdx.x + 1], s2);
        const unsigned ticket = __hip_atomic_fetch_add(&st->ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = ticket == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    double p = 0, d = 0, c = 0;
    for (int b = threadIdx.x; b < a.primal_blocks; b += kBlock) p += load_shared(&a.primal_partial[b]);
    for (int b = threadIdx.x; b < (int)gridDim.x; b += kBlock) {
        d += load_shared(&a.partial[2 * b]);
        c += load_shared(&a.partial[2 * b + 1]);
    }
    p = block_sum(p, lds);
    d = block_sum(d, lds);
    c = block_sum(c, lds);
    if (threadIdx.x == 0) {
        decide(st, a.f, p, d, c);
        __hip_atomic_store(&st->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <int G, class O, class V>
__global__ __launch_bounds__(kBlock) void spmv_kernel(DevCsr<O, V> A, const double *x, double *out) {
    constexpr int rows_per_block = kBlock / G;
    const int lane = threadIdx.x % G;
    for (long long base = (long long)blockIdx.x * rows_per_block; base < A.rows;
         base += (long long)gridDim.x * rows_per_block) {
        const long long i = base + threadIdx.x / G;
        const bool valid = i < A.rows;
        double acc = 0;
        if (valid) {
            const O end = A.ptr[i + 1];
            for (O k = A.ptr[i] + lane; k < end; k += G) acc += double(A.val[k]) * x[A.idx[k]];
        }
        acc = group_sum<G>(acc);
        if (valid && lane == 0) out[i] = acc;
    }
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

__global__ void fill_kernel(double *p, long long n, double v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void copy_kernel(double *dst, const double *src, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
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
    Backend(const npdhg_lp &lp, const double *x0, const double *y0, const npdhg_config &cfg) {
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
        const bool fits32 = nnz_ < (1LL << 31);
        if (cfg.index_bits == 32 && !fits32) throw std::runtime_error("CSR offsets do not fit 32 bits");
        idx64_ = cfg.index_bits == 64 || !fits32;
        HIP_OK(hipGetDevice(&device_));
        HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        try {
            upload_matrix(lp.rows, lp.cols, lp.ptr, lp.index, lp.value, a_);
            HostCsr t = transpose(lp);
            upload_matrix(t.rows, t.cols, t.ptr.data(), t.index.data(), t.value.data(), at_);
            g_a_ = pick_group(m_, nnz_);
            g_at_ = pick_group(n_, nnz_);
            nb1_ = grid_for(n_, g_at_);
            nb2_ = grid_for(m_, g_a_);
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
            if (flags_.halpern) {
                anchor_x_ = alloc_d(n_);
                anchor_y_ = alloc_d(m_);
                anchor_ax_ = alloc_d(m_);
            }
            part1_ = alloc_d(nb1_);
            part2_ = alloc_d(2LL * nb2_);
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
        d.rows = rows;
        d.cols = cols;
        d.nnz = ptr[rows];
        for (long long k = 0; k < d.nnz; ++k)
            if (index[k] < 0 || index[k] >= cols) throw std::runtime_error("CSR column index out of range");
        for (long long i = 0; i < rows; ++i)
            if (ptr[i] > ptr[i + 1] || ptr[i] < 0) throw std::runtime_error("CSR offsets not monotone");
        if (ptr[0] != 0) throw std::runtime_error("CSR offsets must start at 0");
        if (idx64_) {
            d.ptr = alloc_bytes(size_t(rows + 1) * 8);
            HIP_OK(hipMemcpyAsync(d.ptr, ptr, size_t(rows + 1) * 8, hipMemcpyHostToDevice, stream_));
        } else {
            std::vector<int32_t> p32(ptr, ptr + rows + 1);
            d.ptr = alloc_bytes(size_t(rows + 1) * 4);
            HIP_OK(hipMemcpy(d.ptr, p32.data(), p32.size() * 4, hipMemcpyHostToDevice));
        }
        {
            std::vector<int32_t> i32(index, index + d.nnz);
            d.idx = static_cast<int *>(alloc_bytes(size_t(d.nnz) * 4));
            if (d.nnz) HIP_OK(hipMemcpy(d.idx, i32.data(), i32.size() * 4, hipMemcpyHostToDevice));
        }
        if (fp32_) {
            std::vector<float> v32(value, value + d.nnz);
            d.val = alloc_bytes(size_t(d.nnz) * 4);
            if (d.nnz) HIP_OK(hipMemcpy(d.val, v32.data(), v32.size() * 4, hipMemcpyHostToDevice));
        } else {
            d.val = alloc_bytes(size_t(d.nnz) * 8);
            if (d.nnz) HIP_OK(hipMemcpyAsync(d.val, value, size_t(d.nnz) * 8, hipMemcpyHostToDevice, stream_));
        }
    }

    template <class O, class V>
    DevCsr<O, V> view(const DeviceMatrix &d) const {
        return DevCsr<O, V>{static_cast<const O *>(d.ptr), d.idx, static_cast<const V *>(d.val), d.rows};
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
        }
        p.dx = dx_;
        p.anchor_x = anchor_x_;
        p.partial = part1_;
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
        }
        d.dx = dx_;
        d.anchor_y = anchor_y_;
        d.anchor_ax = anchor_ax_;
        d.partial = part2_;
        d.primal_partial = part1_;
        d.primal_blocks = nb1_;
        d.st = st_;
        d.f = flags_;
        return d;
    }

    template <class O, class V>
    void launch_attempt_t() {
        auto at = view<O, V>(at_);
        auto a = view<O, V>(a_);
        auto pa = primal_args();
        auto da = dual_args();
        switch (g_at_) {
#define PRIMAL_CASE(G) \
    case G: primal_kernel<G, O, V><<<nb1_, kBlock, 0, stream_>>>(at, pa); break;
            PRIMAL_CASE(1) PRIMAL_CASE(2) PRIMAL_CASE(4) PRIMAL_CASE(8) PRIMAL_CASE(16) PRIMAL_CASE(32) PRIMAL_CASE(64)
#undef PRIMAL_CASE
        }
        switch (g_a_) {
#define DUAL_CASE(G) \
    case G: dual_kernel<G, O, V><<<nb2_, kBlock, 0, stream_>>>(a, da); break;
            DUAL_CASE(1) DUAL_CASE(2) DUAL_CASE(4) DUAL_CASE(8) DUAL_CASE(16) DUAL_CASE(32) DUAL_CASE(64)
#undef DUAL_CASE
        }
    }
    void launch_attempt() {
        if (idx64_) {
            if (fp32_) launch_attempt_t<int64_t, float>(); else launch_attempt_t<int64_t, double>();
        } else {
            if (fp32_) launch_attempt_t<int32_t, float>(); else launch_attempt_t<int32_t, double>();
        }
        HIP_OK(hipGetLastError());
    }

    template <class O, class V>
    void spmv_t(bool transpose, const double *x, double *out) {
        const DeviceMatrix &d = transpose ? at_ : a_;
        auto mat = view<O, V>(d);
        const int g = transpose ? g_at_ : g_a_;
        const int grid = grid_for(d.rows, g);
        switch (g) {
#define SPMV_CASE(G) \
    case G: spmv_kernel<G, O, V><<<grid, kBlock, 0, stream_>>>(mat, x, out); break;
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
        HIP_OK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
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

    void sync_state() {
        HIP_OK(hipMemcpyAsync(&host_, st_, sizeof(State), hipMemcpyDeviceToHost, stream_));
        HIP_OK(hipStreamSynchronize(stream_));
    }

    void set_point(const double *x0, const double *y0) {
        if (n_) HIP_OK(hipMemcpyAsync(x_[0], x0, n_ * sizeof(double), hipMemcpyHostToDevice, stream_));
        if (m_) HIP_OK(hipMemcpyAsync(y_[0], y0, m_ * sizeof(double), hipMemcpyHostToDevice, stream_));
        spmv(false, x_[0], ax_[0]);
        if (flags_.halpern) {
            copy(anchor_x_, x_[0], n_);
            copy(anchor_y_, y_[0], m_);
            copy(anchor_ax_, ax_[0], m_);
        }
    }

    // x_[c], y_[c] hold the restart point; rebuild activity, clear averages.
    void restart_resident(int c) {
        spmv(false, x_[c], ax_[c]);
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
    int g_a_ = 1, g_at_ = 1, nb1_ = 1, nb2_ = 1;
    Flags flags_{};
    DeviceMatrix a_, at_;
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
    npdhg_backend(const npdhg_lp &lp, const double *x0, const double *y0, const npdhg_config &cfg)
        : impl(lp, x0, y0, cfg) {}
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

int npdhg_create(const npdhg_lp *lp, const double *x0, const double *y0, const npdhg_config *cfg,
                 npdhg_backend **out) {
    return guarded([&] {
        if (!lp || !x0 || !y0 || !cfg || !out) throw std::runtime_error("Null argument");
        *out = new npdhg_backend(*lp, x0, y0, *cfg);
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

int npdhg_spmv(npdhg_backend *b, int transpose, const double *x, double *y) {
    return guarded([&] { b->impl.spmv_host(transpose != 0, x, y); });
}

}  // extern "C"
