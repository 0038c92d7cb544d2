// Device-side Ruiz scaling, objective scaling and CSR transpose of a
// presolved LP (include/niryukti_prepare.h).
//
// The host prepare step (src/preprocess.cpp:133-346) keeps presolve, whose
// guards need long double.  What follows presolve is pure per-row and
// per-column arithmetic over the matrix, so it runs here:
//
//   transpose structure  column counts (atomics) -> exclusive scan -> slots by
//                        atomics -> each entry's rank among its column's
//                        entries (entry ids grow with the row) -> stable order
//   Ruiz pass            one wave per row: max |a|, factor, safety test over
//                        the row (src/preprocess.cpp:12-35), apply; then one
//                        wave per column through the transpose order
//                        (src/preprocess.cpp:36-65)
//   objective scale      max |c| (one workgroup), c /= it
//   gather               A^T index / value in the stable order
//
// Every factor is a max (order-free) followed by the same IEEE FP64 sqrt,
// divide, compare and multiply as the host, so the result is bit-identical to
// the host passes.  The rank step costs sum over columns of degree^2 loads;
// npdhg_scale_lp refuses matrices where that exceeds kRankBudget.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "niryukti_prepare.h"

namespace {

thread_local std::string g_error;

#define HIP_OK(expr)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kScanItems = 8;
constexpr double kRankBudget = 4e10;  // sum of squared column degrees

// A scaling step may never erase a nonzero or turn a finite value infinite
// (src/preprocess.cpp:6-9).
__device__ __forceinline__ bool keeps(double before, double after) {
    return isfinite(before) ? (isfinite(after) && (before == 0 || after != 0)) : before == after;
}

// src/preprocess.cpp:311-313, 319-321: 1/sqrt(max |a|) clipped to
// [1e-3, 1e3] (std::clamp's comparisons); 1 for an empty row or column.
__device__ __forceinline__ double ruiz_factor(double norm) {
    if (!(norm > 0)) return 1.0;
    const double f = 1.0 / sqrt(norm);
    return f < 1e-3 ? 1e-3 : (1e3 < f ? 1e3 : f);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double w = __shfl_xor(v, off, 64);
        v = v < w ? w : v;
    }
    return v;
}

__device__ __forceinline__ long long wave_index() {
    return (long long)blockIdx.x * kWaves + (threadIdx.x >> 6);
}

// --------------------------------------------------------------- transpose --

__global__ void count_columns_kernel(const int64_t *idx, long long nnz, unsigned long long *count) {
    for (long long k = (long long)blockIdx.x * kBlock + threadIdx.x; k < nnz; k += (long long)gridDim.x * kBlock)
        atomicAdd(&count[idx[k]], 1ULL);
}

// Exclusive scan, level 1: each workgroup scans kBlock * kScanItems values
// and leaves its total in sums[blockIdx.x].
__global__ void scan_blocks_kernel(const unsigned long long *in, int64_t *out, long long n, int64_t *sums) {
    __shared__ int64_t wave_total[kWaves];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long long base = ((long long)blockIdx.x * kBlock + threadIdx.x) * kScanItems;
    int64_t v[kScanItems], own = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        v[q] = base + q < n ? int64_t(in[base + q]) : 0;
        own += v[q];
    }
    int64_t inc = own;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int64_t t = __shfl_up(inc, off, 64);
        if (lane >= off) inc += t;
    }
    if (lane == 63) wave_total[wave] = inc;
    __syncthreads();
    int64_t before = 0;
    for (int w = 0; w < wave; ++w) before += wave_total[w];
    int64_t run = before + inc - own;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        if (base + q < n) out[base + q] = run;
        run += v[q];
    }
    if (threadIdx.x == kBlock - 1) sums[blockIdx.x] = run;
}

// Level 2: one workgroup turns the block totals into exclusive offsets.
__global__ void scan_sums_kernel(int64_t *sums, long long count) {
    __shared__ int64_t wave_total[kWaves];
    __shared__ int64_t carry;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (long long c0 = 0; c0 < count; c0 += kBlock) {
        const long long i = c0 + threadIdx.x;
        const int64_t own = i < count ? sums[i] : 0;
        int64_t inc = own;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int64_t t = __shfl_up(inc, off, 64);
            if (lane >= off) inc += t;
        }
        if (lane == 63) wave_total[wave] = inc;
        __syncthreads();
        int64_t before = carry;
        for (int w = 0; w < wave; ++w) before += wave_total[w];
        const int64_t excl = before + inc - own;
        __syncthreads();  // every thread has read carry and wave_total
        if (i < count) sums[i] = excl;
        if (threadIdx.x == kBlock - 1) carry = excl + own;
        __syncthreads();
    }
}

// Level 3: add each workgroup's offset; the last offset is the total.
__global__ void scan_add_kernel(int64_t *out, long long n, const int64_t *sums, int64_t total) {
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i <= n; i += (long long)gridDim.x * kBlock)
        out[i] = i == n ? total : out[i] + sums[i / (kBlock * kScanItems)];
}

// Each entry takes a slot in its column (in no particular order).
__global__ void place_kernel(const int64_t *idx, long long nnz, const int64_t *tptr, unsigned long long *fill,
                             int64_t *slots) {
    for (long long k = (long long)blockIdx.x * kBlock + threadIdx.x; k < nnz; k += (long long)gridDim.x * kBlock) {
        const int64_t j = idx[k];
        const unsigned long long s = atomicAdd(&fill[j], 1ULL);
        slots[tptr[j] + int64_t(s)] = k;
    }
}

// Entry ids grow with the row (CSR order), so an entry's place in its column
// is the number of the column's entries with a smaller id: that order is the
// row order Sparse::transpose produces (src/model.cpp:39-58).
__global__ void rank_kernel(const int64_t *idx, long long nnz, const int64_t *tptr, const int64_t *slots,
                            int64_t *perm) {
    for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < nnz; t += (long long)gridDim.x * kBlock) {
        const int64_t e = slots[t];
        const int64_t j = idx[e], b = tptr[j], end = tptr[j + 1];
        int64_t r = 0;
        for (int64_t s = b; s < end; ++s) r += slots[s] < e;
        perm[b + r] = e;
    }
}

__global__ void row_of_kernel(const int64_t *ptr, long long m, int64_t *row) {
    const long long i = wave_index();
    if (i >= m) return;
    for (int64_t k = ptr[i] + (threadIdx.x & 63); k < ptr[i + 1]; k += 64) row[k] = i;
}

__global__ void gather_transpose_kernel(const int64_t *perm, long long nnz, const int64_t *row, const double *val,
                                        int64_t *t_index, double *t_value) {
    for (long long t = (long long)blockIdx.x * kBlock + threadIdx.x; t < nnz; t += (long long)gridDim.x * kBlock) {
        const int64_t e = perm[t];
        t_index[t] = row[e];
        t_value[t] = val[e];
    }
}

// -------------------------------------------------------------- Ruiz pass --

// Row half of a pass (src/preprocess.cpp:308-314 factors, :18-35 applied):
// a row whose factor would erase or overflow any of its coefficients, its
// bounds or its accumulated scale keeps factor 1.
__global__ void ruiz_rows_kernel(const int64_t *ptr, double *val, long long m, double *rl, double *ru,
                                 double *row_scale) {
    const long long i = wave_index();
    if (i >= m) return;  // the whole wave
    const int lane = threadIdx.x & 63;
    const int64_t b = ptr[i], e = ptr[i + 1];
    double norm = 0;
    for (int64_t k = b + lane; k < e; k += 64) {
        const double a = fabs(val[k]);
        norm = norm < a ? a : norm;
    }
    const double f = ruiz_factor(wave_max(norm));
    if (f == 1.0) return;
    bool ok = true;
    for (int64_t k = b + lane; k < e; k += 64) ok = ok && keeps(val[k], val[k] * f);
    ok = ok && keeps(row_scale[i], row_scale[i] * f) && keeps(rl[i], rl[i] * f) && keeps(ru[i], ru[i] * f);
    if (!__all(ok)) return;
    for (int64_t k = b + lane; k < e; k += 64) val[k] *= f;
    if (lane == 0) {
        row_scale[i] *= f;
        rl[i] *= f;
        ru[i] *= f;
    }
}

// Column half (src/preprocess.cpp:316-323 factors, :36-65 applied): the
// column's own tests first (scale, cost, bounds), then every coefficient.
__global__ void ruiz_columns_kernel(const int64_t *tptr, const int64_t *perm, double *val, long long n, double *c,
                                    double *lb, double *ub, double *column_scale) {
    const long long j = wave_index();
    if (j >= n) return;
    const int lane = threadIdx.x & 63;
    const int64_t b = tptr[j], e = tptr[j + 1];
    double norm = 0;
    for (int64_t t = b + lane; t < e; t += 64) {
        const double a = fabs(val[perm[t]]);
        norm = norm < a ? a : norm;
    }
    const double f = ruiz_factor(wave_max(norm));
    if (f == 1.0) return;
    const double s = column_scale[j], cj = c[j], l = lb[j], u = ub[j];
    if (!keeps(s, s * f) || !keeps(cj, cj * f) || !keeps(0.0, 0.0 * (f * f)) || !keeps(l, l / f) ||
        !keeps(u, u / f))
        return;
    bool ok = true;
    for (int64_t t = b + lane; t < e; t += 64) {
        const double a = val[perm[t]];
        ok = ok && keeps(a, a * f);
    }
    if (!__all(ok)) return;
    for (int64_t t = b + lane; t < e; t += 64) val[perm[t]] *= f;
    if (lane == 0) {
        column_scale[j] = s * f;
        c[j] = cj * f;
        lb[j] = l / f;
        ub[j] = u / f;
    }
}

// Objective scale (src/preprocess.cpp:336-345): max |c|, 1 when it is 0.
// Level 1: each workgroup's max; level 2 (one workgroup): the max of those.
__global__ void abs_max_kernel(const double *c, long long n, double *part) {
    __shared__ double wave_part[kWaves];
    double v = 0;
    for (long long j = (long long)blockIdx.x * kBlock + threadIdx.x; j < n; j += (long long)gridDim.x * kBlock) {
        const double a = fabs(c[j]);
        v = v < a ? a : v;
    }
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) wave_part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0;
        for (int w = 0; w < kWaves; ++w) m = m < wave_part[w] ? wave_part[w] : m;
        part[blockIdx.x] = m;
    }
}

__global__ void objective_scale_kernel(const double *part, int count, double *scale) {
    __shared__ double wave_part[kWaves];
    double v = 0;
    for (int j = threadIdx.x; j < count; j += kBlock) v = v < part[j] ? part[j] : v;
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) wave_part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0;
        for (int w = 0; w < kWaves; ++w) m = m < wave_part[w] ? wave_part[w] : m;
        *scale = m > 0 ? m : 1.0;
    }
}

__global__ void divide_kernel(double *c, long long n, const double *scale) {
    const double s = *scale;
    for (long long j = (long long)blockIdx.x * kBlock + threadIdx.x; j < n; j += (long long)gridDim.x * kBlock)
        c[j] /= s;
}

// ------------------------------------------------------------------- host --

int grid_for(long long items) {
    return int(std::max<long long>(1, std::min<long long>((items + kBlock - 1) / kBlock, 65536)));
}
int wave_grid(long long rows) { return int(std::max<long long>(1, (rows + kWaves - 1) / kWaves)); }

class DeviceBuffers {
  public:
    ~DeviceBuffers() {
        for (void *p : held_) (void)hipFree(p);
    }
    template <class T>
    T *alloc(long long count) {
        void *p = nullptr;
        HIP_OK(hipMalloc(&p, size_t(std::max<long long>(count, 1)) * sizeof(T)));
        held_.push_back(p);
        return static_cast<T *>(p);
    }

  private:
    std::vector<void *> held_;
};

void scale_lp(const npdhg_lp &lp, int passes, npdhg_scaled &out) {
    using Clock = std::chrono::steady_clock;
    const auto t0 = Clock::now();
    const long long m = lp.rows, n = lp.cols;
    if (m < 0 || n < 0 || m >= (1LL << 31) || n >= (1LL << 31)) throw std::runtime_error("LP dimensions out of range");
    if (passes < 0) throw std::runtime_error("Negative scaling passes");
    if (!lp.ptr || (n && (!lp.c || !lp.lb || !lp.ub)) || (m && (!lp.rl || !lp.ru)))
        throw std::runtime_error("Null LP array");
    if (lp.ptr[0] != 0) throw std::runtime_error("CSR offsets must start at 0");
    const long long nnz = lp.ptr[m];
    if (nnz >= (1LL << 31)) throw std::runtime_error("Device scaling needs fewer than 2^31 nonzeros");
    for (long long i = 0; i < m; ++i)
        if (lp.ptr[i] > lp.ptr[i + 1]) throw std::runtime_error("CSR offsets not monotone");
    if (nnz && (!lp.index || !lp.value)) throw std::runtime_error("Null LP array");
    if (!out.value || (n && (!out.c || !out.lb || !out.ub || !out.column_scale)) ||
        (m && (!out.rl || !out.ru || !out.row_scale)) || !out.t_ptr || (nnz && (!out.t_index || !out.t_value)))
        throw std::runtime_error("Null output array");
    // Column degrees on the host: range check and the rank step's cost.
    {
        std::vector<int64_t> degree(size_t(n), 0);
        for (long long k = 0; k < nnz; ++k) {
            const int64_t j = lp.index[k];
            if (j < 0 || j >= n) throw std::runtime_error("CSR column index out of range");
            if (!std::isfinite(lp.value[k])) throw std::runtime_error("Non-finite coefficient");
            ++degree[size_t(j)];
        }
        double work = 0;
        for (int64_t d : degree) work += double(d) * double(d);
        if (work > kRankBudget) throw std::runtime_error("Columns too long for the device transpose");
    }

    DeviceBuffers d;  // freed after the guard below has synchronised the stream
    hipStream_t s;
    HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t ev[2] = {nullptr, nullptr};
    struct StreamGuard {
        hipStream_t s;
        hipEvent_t *ev;
        ~StreamGuard() {
            (void)hipStreamSynchronize(s);
            for (int k = 0; k < 2; ++k)
                if (ev[k]) (void)hipEventDestroy(ev[k]);
            (void)hipStreamDestroy(s);
        }
    } guard{s, ev};
    HIP_OK(hipEventCreate(&ev[0]));
    HIP_OK(hipEventCreate(&ev[1]));

    int64_t *ptr = d.alloc<int64_t>(m + 1), *idx = d.alloc<int64_t>(nnz);
    double *val = d.alloc<double>(nnz);
    double *c = d.alloc<double>(n), *lb = d.alloc<double>(n), *ub = d.alloc<double>(n);
    double *rl = d.alloc<double>(m), *ru = d.alloc<double>(m);
    double *rs = d.alloc<double>(m), *cs = d.alloc<double>(n), *osc = d.alloc<double>(1);
    const int max_blocks = grid_for(n);
    double *cmax = d.alloc<double>(max_blocks);
    unsigned long long *count = d.alloc<unsigned long long>(n);
    int64_t *tptr = d.alloc<int64_t>(n + 1);
    const long long nsums = std::max<long long>(1, (n + kBlock * kScanItems - 1) / (kBlock * kScanItems));
    int64_t *sums = d.alloc<int64_t>(nsums);
    int64_t *slots = d.alloc<int64_t>(nnz), *perm = d.alloc<int64_t>(nnz), *row = d.alloc<int64_t>(nnz);
    int64_t *t_index = d.alloc<int64_t>(nnz);
    double *t_value = d.alloc<double>(nnz);

    auto up = [&](void *dst, const void *src, size_t bytes) {
        if (bytes) HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    };
    up(ptr, lp.ptr, size_t(m + 1) * 8);
    up(idx, lp.index, size_t(nnz) * 8);
    up(val, lp.value, size_t(nnz) * 8);
    up(c, lp.c, size_t(n) * 8);
    up(lb, lp.lb, size_t(n) * 8);
    up(ub, lp.ub, size_t(n) * 8);
    up(rl, lp.rl, size_t(m) * 8);
    up(ru, lp.ru, size_t(m) * 8);
    const std::vector<double> ones(size_t(std::max(m, n)), 1.0);
    up(rs, ones.data(), size_t(m) * 8);
    up(cs, ones.data(), size_t(n) * 8);
    HIP_OK(hipMemsetAsync(count, 0, size_t(std::max<long long>(n, 1)) * 8, s));
    HIP_OK(hipEventRecord(ev[0], s));

    // transpose structure
    if (nnz) count_columns_kernel<<<grid_for(nnz), kBlock, 0, s>>>(idx, nnz, count);
    scan_blocks_kernel<<<int(nsums), kBlock, 0, s>>>(count, tptr, n, sums);
    scan_sums_kernel<<<1, kBlock, 0, s>>>(sums, nsums);
    scan_add_kernel<<<grid_for(n + 1), kBlock, 0, s>>>(tptr, n, sums, int64_t(nnz));
    HIP_OK(hipMemsetAsync(count, 0, size_t(std::max<long long>(n, 1)) * 8, s));
    if (nnz) {
        place_kernel<<<grid_for(nnz), kBlock, 0, s>>>(idx, nnz, tptr, count, slots);
        rank_kernel<<<grid_for(nnz), kBlock, 0, s>>>(idx, nnz, tptr, slots, perm);
    }
    if (m) row_of_kernel<<<wave_grid(m), kBlock, 0, s>>>(ptr, m, row);
    HIP_OK(hipGetLastError());

    // Ruiz passes (src/preprocess.cpp:333-334), then the objective scale
    for (int pass = 0; pass < passes; ++pass) {
        if (m) ruiz_rows_kernel<<<wave_grid(m), kBlock, 0, s>>>(ptr, val, m, rl, ru, rs);
        if (n) ruiz_columns_kernel<<<wave_grid(n), kBlock, 0, s>>>(tptr, perm, val, n, c, lb, ub, cs);
    }
    if (passes > 0) {
        abs_max_kernel<<<max_blocks, kBlock, 0, s>>>(c, n, cmax);
        objective_scale_kernel<<<1, kBlock, 0, s>>>(cmax, max_blocks, osc);
        if (n) divide_kernel<<<grid_for(n), kBlock, 0, s>>>(c, n, osc);
    }
    if (nnz) gather_transpose_kernel<<<grid_for(nnz), kBlock, 0, s>>>(perm, nnz, row, val, t_index, t_value);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(ev[1], s));

    auto down = [&](void *dst, const void *src, size_t bytes) {
        if (bytes) HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    };
    down(out.value, val, size_t(nnz) * 8);
    down(out.c, c, size_t(n) * 8);
    down(out.lb, lb, size_t(n) * 8);
    down(out.ub, ub, size_t(n) * 8);
    down(out.rl, rl, size_t(m) * 8);
    down(out.ru, ru, size_t(m) * 8);
    down(out.row_scale, rs, size_t(m) * 8);
    down(out.column_scale, cs, size_t(n) * 8);
    down(out.t_ptr, tptr, size_t(n + 1) * 8);
    down(out.t_index, t_index, size_t(nnz) * 8);
    down(out.t_value, t_value, size_t(nnz) * 8);
    double scale = 1.0;
    if (passes > 0) down(&scale, osc, 8);
    HIP_OK(hipStreamSynchronize(s));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, ev[0], ev[1]));
    out.objective_scale = scale;
    out.device_seconds = ms * 1e-3;
    out.seconds = std::chrono::duration<double>(Clock::now() - t0).count();
}

// ------------------------------------------------ no-allocation variant --
//
// npdhg_scale_lp_noalloc runs the same kernels on LP arrays that are already
// in device memory and writes into caller-provided device arrays; its work
// arrays (column counts, slots, ranks, row of each entry, scan sums, the
// objective-scale partials and the validation flags) live in one static
// device array, `scratch`, so it never allocates device memory.

constexpr size_t kScratchBytes = size_t(128) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Per-device address of `scratch` (each device holds its own copy of the
// array); filled on the first call on a device with that device current.
constexpr int kMaxDevices = 64;
std::atomic<unsigned char *> g_scratch[kMaxDevices];

unsigned char *scratch_address() {
    int device = 0;
    HIP_OK(hipGetDevice(&device));
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("Device ordinal out of range for the scratch table");
    unsigned char *p = g_scratch[device].load(std::memory_order_acquire);
    if (!p) {
        void *addr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !addr)
            throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        p = static_cast<unsigned char *>(addr);
        g_scratch[device].store(p, std::memory_order_release);
    }
    return p;
}

__global__ void fill_u64_kernel(unsigned long long *p, long long n, unsigned long long v) {
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock) p[i] = v;
}
__global__ void fill_f64_kernel(double *p, long long n, double v) {
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock) p[i] = v;
}

// flags[0]: a row's offsets decrease or leave [0, nnz]; flags[1]: a column
// index outside [0, n); flags[2]: a non-finite coefficient.
__global__ void validate_kernel(const int64_t *ptr, long long m, const int64_t *idx, const double *val, long long nnz,
                                long long n, unsigned int *flags) {
    const long long stride = (long long)gridDim.x * kBlock;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < m; i += stride)
        if (ptr[i] > ptr[i + 1] || ptr[i] < 0 || ptr[i + 1] > nnz) atomicOr(&flags[0], 1u);
    for (long long k = (long long)blockIdx.x * kBlock + threadIdx.x; k < nnz; k += stride) {
        const int64_t j = idx[k];
        if (j < 0 || j >= n) atomicOr(&flags[1], 1u);
        if (!isfinite(val[k])) atomicOr(&flags[2], 1u);
    }
}

// Sum of squared column degrees (the rank step's cost), one atomic per block.
__global__ void degree_work_kernel(const unsigned long long *count, long long n, unsigned long long *work) {
    __shared__ unsigned long long part[kBlock];
    unsigned long long w = 0;
    for (long long j = (long long)blockIdx.x * kBlock + threadIdx.x; j < n; j += (long long)gridDim.x * kBlock)
        w += count[j] * count[j];
    part[threadIdx.x] = w;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h /= 2) {
        if (int(threadIdx.x) < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(work, part[0]);
}

// Byte layout of the work arrays in `scratch` for an LP of n columns and nnz
// nonzeros; total() beyond kScratchBytes is refused.
struct ScratchPlan {
    size_t flags, work, osc, cmax, count, sums, slots, perm, row, end;
    ScratchPlan(long long n, long long nnz, int cmax_blocks, long long nsums) {
        size_t at = 0;
        auto take = [&](size_t bytes) {
            const size_t here = at;
            at = (at + std::max<size_t>(bytes, 8) + 255) & ~size_t(255);
            return here;
        };
        flags = take(4 * sizeof(unsigned int));
        work = take(sizeof(unsigned long long));
        osc = take(sizeof(double));
        cmax = take(size_t(cmax_blocks) * sizeof(double));
        count = take(size_t(n) * sizeof(unsigned long long));
        sums = take(size_t(nsums) * sizeof(int64_t));
        slots = take(size_t(nnz) * sizeof(int64_t));
        perm = take(size_t(nnz) * sizeof(int64_t));
        row = take(size_t(nnz) * sizeof(int64_t));
        end = at;
    }
};

void scale_lp_noalloc(const npdhg_lp &lp, int passes, npdhg_scaled &out, hipStream_t s) {
    const long long m = lp.rows, n = lp.cols;
    if (m < 0 || n < 0 || m >= (1LL << 31) || n >= (1LL << 31)) throw std::runtime_error("LP dimensions out of range");
    if (passes < 0) throw std::runtime_error("Negative scaling passes");
    if (!lp.ptr || (n && (!lp.c || !lp.lb || !lp.ub)) || (m && (!lp.rl || !lp.ru)))
        throw std::runtime_error("Null LP array");
    if (!out.value || (n && (!out.c || !out.lb || !out.ub || !out.column_scale)) ||
        (m && (!out.rl || !out.ru || !out.row_scale)) || !out.t_ptr)
        throw std::runtime_error("Null output array");
    unsigned char *base = scratch_address();
    // the matrix's extent, from its device offsets
    int64_t ends[2] = {0, 0};
    HIP_OK(hipMemcpyAsync(&ends[0], lp.ptr, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(&ends[1], lp.ptr + m, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (ends[0] != 0) throw std::runtime_error("CSR offsets must start at 0");
    const long long nnz = ends[1];
    if (nnz < 0 || nnz >= (1LL << 31)) throw std::runtime_error("Device scaling needs 0 <= nonzeros < 2^31");
    if (nnz && (!lp.index || !lp.value || !out.t_index || !out.t_value)) throw std::runtime_error("Null matrix array");
    const int cmax_blocks = grid_for(n);
    const long long nsums = std::max<long long>(1, (n + kBlock * kScanItems - 1) / (kBlock * kScanItems));
    const ScratchPlan plan(n, nnz, cmax_blocks, nsums);
    if (plan.end > kScratchBytes)
        throw std::runtime_error("LP needs " + std::to_string(plan.end) + " bytes of scratch; the no-allocation entry "
                                 "point holds " + std::to_string(kScratchBytes) + " (24 bytes per nonzero + 8 per column)");
    auto *flags = reinterpret_cast<unsigned int *>(base + plan.flags);
    auto *work = reinterpret_cast<unsigned long long *>(base + plan.work);
    auto *osc = reinterpret_cast<double *>(base + plan.osc);
    auto *cmax = reinterpret_cast<double *>(base + plan.cmax);
    auto *count = reinterpret_cast<unsigned long long *>(base + plan.count);
    auto *sums = reinterpret_cast<int64_t *>(base + plan.sums);
    auto *slots = reinterpret_cast<int64_t *>(base + plan.slots);
    auto *perm = reinterpret_cast<int64_t *>(base + plan.perm);
    auto *row = reinterpret_cast<int64_t *>(base + plan.row);

    // flags and the degree sum start at zero on every call (fill kernels:
    // the scratch keeps the previous call's contents)
    fill_u64_kernel<<<1, kBlock, 0, s>>>(reinterpret_cast<unsigned long long *>(flags), 2, 0);
    fill_u64_kernel<<<1, kBlock, 0, s>>>(work, 1, 0);
    validate_kernel<<<grid_for(std::max(m, nnz)), kBlock, 0, s>>>(lp.ptr, m, lp.index, lp.value, nnz, n, flags);
    HIP_OK(hipGetLastError());
    unsigned int seen[3] = {0, 0, 0};
    HIP_OK(hipMemcpyAsync(seen, flags, sizeof seen, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (seen[0]) throw std::runtime_error("CSR offsets not monotone");
    if (seen[1]) throw std::runtime_error("CSR column index out of range");
    if (seen[2]) throw std::runtime_error("Non-finite coefficient");

    // outputs start as copies of the inputs and are scaled in place
    auto copy = [&](double *dst, const double *src, long long count_) {
        if (count_) HIP_OK(hipMemcpyAsync(dst, src, size_t(count_) * 8, hipMemcpyDeviceToDevice, s));
    };
    copy(out.value, lp.value, nnz);
    copy(out.c, lp.c, n);
    copy(out.lb, lp.lb, n);
    copy(out.ub, lp.ub, n);
    copy(out.rl, lp.rl, m);
    copy(out.ru, lp.ru, m);
    if (m) fill_f64_kernel<<<grid_for(m), kBlock, 0, s>>>(out.row_scale, m, 1.0);
    if (n) fill_f64_kernel<<<grid_for(n), kBlock, 0, s>>>(out.column_scale, n, 1.0);
    if (n) fill_u64_kernel<<<grid_for(n), kBlock, 0, s>>>(count, n, 0);
    if (nnz) count_columns_kernel<<<grid_for(nnz), kBlock, 0, s>>>(lp.index, nnz, count);
    if (n) degree_work_kernel<<<grid_for(n), kBlock, 0, s>>>(count, n, work);
    HIP_OK(hipGetLastError());
    unsigned long long rank_work = 0;
    HIP_OK(hipMemcpyAsync(&rank_work, work, sizeof rank_work, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (double(rank_work) > kRankBudget) throw std::runtime_error("Columns too long for the device transpose");

    int64_t *tptr = out.t_ptr;
    double *val = out.value;
    scan_blocks_kernel<<<int(nsums), kBlock, 0, s>>>(count, tptr, n, sums);
    scan_sums_kernel<<<1, kBlock, 0, s>>>(sums, nsums);
    scan_add_kernel<<<grid_for(n + 1), kBlock, 0, s>>>(tptr, n, sums, int64_t(nnz));
    if (n) fill_u64_kernel<<<grid_for(n), kBlock, 0, s>>>(count, n, 0);
    if (nnz) {
        place_kernel<<<grid_for(nnz), kBlock, 0, s>>>(lp.index, nnz, tptr, count, slots);
        rank_kernel<<<grid_for(nnz), kBlock, 0, s>>>(lp.index, nnz, tptr, slots, perm);
    }
    if (m) row_of_kernel<<<wave_grid(m), kBlock, 0, s>>>(lp.ptr, m, row);
    for (int pass = 0; pass < passes; ++pass) {
        if (m) ruiz_rows_kernel<<<wave_grid(m), kBlock, 0, s>>>(lp.ptr, val, m, out.rl, out.ru, out.row_scale);
        if (n)
            ruiz_columns_kernel<<<wave_grid(n), kBlock, 0, s>>>(tptr, perm, val, n, out.c, out.lb, out.ub,
                                                                out.column_scale);
    }
    if (passes > 0) {
        abs_max_kernel<<<cmax_blocks, kBlock, 0, s>>>(out.c, n, cmax);
        objective_scale_kernel<<<1, kBlock, 0, s>>>(cmax, cmax_blocks, osc);
        if (n) divide_kernel<<<grid_for(n), kBlock, 0, s>>>(out.c, n, osc);
    }
    if (nnz) gather_transpose_kernel<<<grid_for(nnz), kBlock, 0, s>>>(perm, nnz, row, val, out.t_index, out.t_value);
    HIP_OK(hipGetLastError());
    double scale = 1.0;
    if (passes > 0) HIP_OK(hipMemcpyAsync(&scale, osc, sizeof scale, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    out.objective_scale = scale;
    out.device_seconds = 0;
    out.seconds = 0;
}

}  // namespace

extern "C" {

const char *npdhg_scale_last_error(void) { return g_error.c_str(); }

int npdhg_scale_lp_noalloc(const npdhg_lp *lp, int passes, npdhg_scaled *out, void *stream) {
    try {
        g_error.clear();
        if (!lp || !out) throw std::runtime_error("Null argument");
        scale_lp_noalloc(*lp, passes, *out, static_cast<hipStream_t>(stream));
        return 0;
    } catch (const std::exception &e) {
        g_error = e.what();
        return -1;
    } catch (...) {
        g_error = "unknown native error";
        return -1;
    }
}

int npdhg_scale_lp(const npdhg_lp *lp, int passes, npdhg_scaled *out) {
    try {
        g_error.clear();
        if (!lp || !out) throw std::runtime_error("Null argument");
        scale_lp(*lp, passes, *out);
        return 0;
    } catch (const std::exception &e) {
        g_error = e.what();
        return -1;
    } catch (...) {
        g_error = "unknown native error";
        return -1;
    }
}

}  // extern "C"
