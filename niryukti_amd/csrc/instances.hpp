// Seeded synthetic LP generators with a planted, certified optimum.
//
// Every instance is built from a planted primal point x0 and a planted dual
// point so that (x0, y) satisfies the KKT conditions exactly (up to the FP64
// rounding of b = A x0 and c = A^T y0 + s):
//
//   * equality rows  : rl = ru = (A x0)_i,                 y0_i ~ N(0,1)
//   * tight <= rows  : rl = -inf, ru = (A x0)_i,           y0_i = -|N(0,1)|
//   * slack <= rows  : rl = -inf, ru = (A x0)_i + U(0,1),  y0_i = 0
//   * columns        : x0_j = 0 with probability zero_fraction, else U(0, upper)
//                      s_j ~ U(0,1) where x0_j = 0, else 0;   c = A^T y0 + s
//
// With the Lagrangian sign convention g = c + A^T y used by the solver
// (src/cpu.cpp:107-112 and src/verify.cpp:177 of the reference), the optimal
// dual is y = -y0: g = s >= 0 is supported only on columns sitting at their
// lower bound 0, and the dual of every <= row is >= 0 (its upper endpoint is
// the active one).  Hence the optimum objective is c^T x0 = b^T y0 exactly.
//
// Randomness comes from std::mt19937_64 only; uniform and normal variates are
// derived here (53-bit uniforms, Marsaglia polar normals) so that instances do
// not depend on the C++ library's distribution implementations.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace nyk_gen {

constexpr double kInf = std::numeric_limits<double>::infinity();

inline uint64_t splitmix(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

struct Rng {
    std::mt19937_64 g;
    bool has_spare = false;
    double spare = 0;
    explicit Rng(uint64_t seed) : g(seed) {}
    double uniform() { return double(g() >> 11) * 0x1.0p-53; }  // [0, 1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    // Unbiased integer in [0, n) by rejection.
    uint64_t below(uint64_t n) {
        uint64_t limit = std::numeric_limits<uint64_t>::max() - std::numeric_limits<uint64_t>::max() % n;
        uint64_t v;
        do v = g(); while (v >= limit);
        return v % n;
    }
    double normal() {
        if (has_spare) {
            has_spare = false;
            return spare;
        }
        double u, v, s;
        do {
            u = 2 * uniform() - 1;
            v = 2 * uniform() - 1;
            s = u * u + v * v;
        } while (s >= 1 || s == 0);
        double f = std::sqrt(-2 * std::log(s) / s);
        spare = v * f;
        has_spare = true;
        return u * f;
    }
};

enum RowKind : int8_t { kEqual = 0, kTight = 1, kSlack = 2 };

struct Instance {
    std::string name;
    int64_t rows = 0, cols = 0;
    std::vector<int64_t> ptr, index;  // CSR of A, column indices ascending per row
    std::vector<double> value;
    std::vector<double> c, lb, ub, rl, ru;
    std::vector<double> x0;     // planted optimal primal point
    std::vector<double> y_opt;  // planted optimal dual in the solver's sign convention (= -y0)
    double objective = 0;       // c^T x0, accumulated in long double
    uint64_t seed = 0;
    int64_t nnz() const { return int64_t(value.size()); }
};

// Column-major triplets (col ascending, rows in draw order) -> CSR with rows
// ascending, columns ascending inside a row, duplicate (row, col) pairs summed
// in draw order and exact zeros dropped.  The counting sort is stable, so a
// row's entries arrive ordered by column already.
inline void build_csr(Instance &p, const std::vector<int64_t> &trow, const std::vector<int64_t> &tcol,
                      const std::vector<double> &tval) {
    const int64_t m = p.rows;
    const size_t t = trow.size();
    std::vector<int64_t> count(m + 1, 0);
    for (size_t k = 0; k < t; ++k) count[trow[k] + 1]++;
    for (int64_t i = 0; i < m; ++i) count[i + 1] += count[i];
    std::vector<int64_t> col(t);
    std::vector<double> val(t);
    {
        std::vector<int64_t> pos(count.begin(), count.end() - 1);
        for (size_t k = 0; k < t; ++k) {
            int64_t q = pos[trow[k]]++;
            col[q] = tcol[k];
            val[q] = tval[k];
        }
    }
    p.ptr.assign(m + 1, 0);
    p.index.clear();
    p.value.clear();
    p.index.reserve(t);
    p.value.reserve(t);
    for (int64_t i = 0; i < m; ++i) {
        int64_t k = count[i], e = count[i + 1];
        while (k < e) {
            int64_t j = col[k];
            double v = val[k++];
            while (k < e && col[k] == j) v += val[k++];
            if (v != 0) {
                p.index.push_back(j);
                p.value.push_back(v);
            }
        }
        p.ptr[i + 1] = int64_t(p.index.size());
    }
}

// b = A x0, c = A^T y0 + s, bounds and planted objective, given row kinds.
inline void plant(Instance &p, const std::vector<int8_t> &kind, const std::vector<double> &y0,
                  const std::vector<double> &s, Rng &rhs_rng) {
    const int64_t m = p.rows, n = p.cols;
    p.rl.assign(m, 0);
    p.ru.assign(m, 0);
    for (int64_t i = 0; i < m; ++i) {
        double b = 0;
        for (int64_t k = p.ptr[i]; k < p.ptr[i + 1]; ++k) b += p.value[k] * p.x0[p.index[k]];
        if (kind[i] == kEqual) {
            p.rl[i] = p.ru[i] = b;
        } else {
            p.rl[i] = -kInf;
            p.ru[i] = kind[i] == kTight ? b : b + rhs_rng.uniform(0.0, 1.0);
        }
    }
    p.c = s;
    for (int64_t i = 0; i < m; ++i)
        for (int64_t k = p.ptr[i]; k < p.ptr[i + 1]; ++k) p.c[p.index[k]] += p.value[k] * y0[i];
    // c was accumulated row by row above; recompute column sums in column
    // order is unnecessary: the planted point is optimal for whatever FP64
    // value c holds, because s and x0 are complementary by construction.
    p.y_opt.resize(m);
    for (int64_t i = 0; i < m; ++i) p.y_opt[i] = -y0[i];
    long double obj = 0;
    for (int64_t j = 0; j < n; ++j) obj += (long double)p.c[j] * p.x0[j];
    p.objective = double(obj);
}

inline void plant_columns(Instance &p, double upper, double zero_fraction, Rng &rng,
                          std::vector<double> &s) {
    const int64_t n = p.cols;
    p.lb.assign(n, 0);
    p.ub.assign(n, upper);
    p.x0.assign(n, 0);
    s.assign(n, 0);
    for (int64_t j = 0; j < n; ++j) {
        if (rng.uniform() < zero_fraction) {
            p.x0[j] = 0;
            s[j] = rng.uniform();
        } else {
            p.x0[j] = rng.uniform(0.0, upper);
        }
    }
}

// Equality-constrained LP, "uniform" structure: every column draws its number
// of nonzeros uniformly from [nnz_lo, nnz_hi] and places them at uniformly
// random rows (with replacement; repeated rows are summed).  Values N(0,1).
// Box [0, upper].  This is the 2,000 x 5,000 CPU workload (5 nnz per column,
// upper 10) and the per-instance generator of the batched / served workloads.
inline Instance equality_lp(int64_t m, int64_t n, int nnz_lo, int nnz_hi, double upper,
                            uint64_t seed, double zero_fraction = 0.3) {
    if (m <= 0 || n <= 0 || nnz_lo < 1 || nnz_hi < nnz_lo || !(upper > 0))
        throw std::invalid_argument("equality_lp: bad shape");
    Instance p;
    p.rows = m;
    p.cols = n;
    p.seed = seed;
    Rng srng(splitmix(seed ^ 0x11)), vrng(splitmix(seed ^ 0x22)), xrng(splitmix(seed ^ 0x33)),
        yrng(splitmix(seed ^ 0x44));
    std::vector<int64_t> trow, tcol;
    std::vector<double> tval;
    trow.reserve(size_t(n) * nnz_hi);
    tcol.reserve(size_t(n) * nnz_hi);
    tval.reserve(size_t(n) * nnz_hi);
    for (int64_t j = 0; j < n; ++j) {
        int d = nnz_lo + int(srng.below(uint64_t(nnz_hi - nnz_lo + 1)));
        for (int t = 0; t < d; ++t) {
            trow.push_back(int64_t(srng.below(uint64_t(m))));
            tcol.push_back(j);
            tval.push_back(vrng.normal());
        }
    }
    build_csr(p, trow, tcol, tval);
    std::vector<double> s;
    plant_columns(p, upper, zero_fraction, xrng, s);
    std::vector<int8_t> kind(m, kEqual);
    std::vector<double> y0(m);
    for (int64_t i = 0; i < m; ++i) y0[i] = yrng.normal();
    plant(p, kind, y0, s, yrng);
    p.name = "equality_" + std::to_string(m) + "x" + std::to_string(n) + "_s" + std::to_string(seed);
    return p;
}

// Power-law LP: 70% equality rows and 30% <= rows (half of those tight, half
// with slack).  Column degrees follow a Zipf law P(d) ~ d^-zipf on
// d = min_degree..max_degree (exponent 1.5 on 2..25 gives a mean of ~6.2
// nonzeros per column, ~12.4M nonzeros at 2,000,000 columns).  Each nonzero
// lands, with probability hub_fraction, in one of `hub_rows` hub rows chosen
// with Zipf(1) popularity (rank r has weight 1/r), otherwise in a uniformly
// random row.  With 64 hubs and 2% of the nonzeros the longest hub row holds
// ~52k nonzeros at full size and the hub lengths decay as 1/r, so row lengths
// are heavy-tailed while the bulk of the rows stays short.  Values N(0,1),
// box [0, upper].
inline Instance powerlaw_lp(int64_t m, int64_t n, uint64_t seed, double zipf = 1.5,
                            int min_degree = 2, int max_degree = 25, int64_t hub_rows = 64,
                            double hub_fraction = 0.02, double upper = 100,
                            double zero_fraction = 0.3, double equality_fraction = 0.7) {
    if (m <= 0 || n <= 0 || min_degree < 1 || max_degree < min_degree || hub_rows < 0)
        throw std::invalid_argument("powerlaw_lp: bad shape");
    hub_rows = std::min(hub_rows, m);
    Instance p;
    p.rows = m;
    p.cols = n;
    p.seed = seed;
    Rng srng(splitmix(seed ^ 0x51)), vrng(splitmix(seed ^ 0x52)), xrng(splitmix(seed ^ 0x53)),
        yrng(splitmix(seed ^ 0x54)), prng(splitmix(seed ^ 0x55));
    std::vector<double> degree_cdf(max_degree - min_degree + 1);
    double acc = 0;
    for (int d = min_degree; d <= max_degree; ++d) degree_cdf[d - min_degree] = acc += std::pow(double(d), -zipf);
    std::vector<double> hub_cdf(hub_rows);
    acc = 0;
    for (int64_t r = 0; r < hub_rows; ++r) hub_cdf[r] = acc += 1.0 / double(r + 1);
    // Hubs sit at uniformly random distinct rows (partial Fisher-Yates).
    std::vector<int64_t> hub(hub_rows);
    {
        std::vector<int64_t> perm(m);
        for (int64_t r = 0; r < m; ++r) perm[r] = r;
        for (int64_t r = 0; r < hub_rows; ++r) {
            int64_t q = r + int64_t(prng.below(uint64_t(m - r)));
            std::swap(perm[r], perm[q]);
            hub[r] = perm[r];
        }
    }
    std::vector<int64_t> trow, tcol;
    std::vector<double> tval;
    size_t guess = size_t(n) * 7;
    trow.reserve(guess);
    tcol.reserve(guess);
    tval.reserve(guess);
    const double dtotal = degree_cdf.back(), htotal = hub_rows ? hub_cdf.back() : 0;
    for (int64_t j = 0; j < n; ++j) {
        double u = srng.uniform() * dtotal;
        int d = int(std::upper_bound(degree_cdf.begin(), degree_cdf.end(), u) - degree_cdf.begin()) + min_degree;
        if (d > max_degree) d = max_degree;
        for (int t = 0; t < d; ++t) {
            int64_t row;
            if (hub_rows && srng.uniform() < hub_fraction) {
                double v = srng.uniform() * htotal;
                int64_t r = int64_t(std::upper_bound(hub_cdf.begin(), hub_cdf.end(), v) - hub_cdf.begin());
                row = hub[std::min(r, hub_rows - 1)];
            } else {
                row = int64_t(srng.below(uint64_t(m)));
            }
            trow.push_back(row);
            tcol.push_back(j);
            tval.push_back(vrng.normal());
        }
    }
    build_csr(p, trow, tcol, tval);
    std::vector<double> s;
    plant_columns(p, upper, zero_fraction, xrng, s);
    std::vector<int8_t> kind(m);
    std::vector<double> y0(m);
    for (int64_t i = 0; i < m; ++i) {
        double u = yrng.uniform();
        kind[i] = u < equality_fraction                               ? kEqual
                  : u < equality_fraction + 0.5 * (1 - equality_fraction) ? kTight
                                                                          : kSlack;
        double g = yrng.normal();
        y0[i] = kind[i] == kEqual ? g : kind[i] == kTight ? -std::abs(g) : 0.0;
    }
    plant(p, kind, y0, s, yrng);
    p.name = "powerlaw_" + std::to_string(m) + "x" + std::to_string(n) + "_s" + std::to_string(seed);
    return p;
}

// Block-angular planning LP: `blocks` diagonal blocks of brows x bcols
// equality rows (nnz_per_col uniform rows inside the block, N(0,1)), plus
// `link_rows` <= linking rows (half tight) that touch a `link_fraction` of all
// columns, each touched column appearing in `link_per_col` uniformly chosen
// linking rows.  Planted duals of block rows are U(-1,1) ("costs U[-1,1]"),
// box [0, upper].  Rows: block rows block by block, then the linking rows.
inline Instance block_angular_lp(int blocks, int64_t brows, int64_t bcols, int nnz_per_col,
                                 int64_t link_rows, double link_fraction, int link_per_col,
                                 uint64_t seed, double upper = 1000, double zero_fraction = 0.3) {
    if (blocks < 1 || brows < 1 || bcols < 1 || nnz_per_col < 1 || link_rows < 0)
        throw std::invalid_argument("block_angular_lp: bad shape");
    Instance p;
    p.rows = int64_t(blocks) * brows + link_rows;
    p.cols = int64_t(blocks) * bcols;
    p.seed = seed;
    Rng srng(splitmix(seed ^ 0x61)), vrng(splitmix(seed ^ 0x62)), xrng(splitmix(seed ^ 0x63)),
        yrng(splitmix(seed ^ 0x64));
    std::vector<int64_t> trow, tcol;
    std::vector<double> tval;
    size_t guess = size_t(p.cols) * nnz_per_col + size_t(double(p.cols) * link_fraction * link_per_col * 1.1);
    trow.reserve(guess);
    tcol.reserve(guess);
    tval.reserve(guess);
    const int64_t link_base = int64_t(blocks) * brows;
    for (int b = 0; b < blocks; ++b)
        for (int64_t jj = 0; jj < bcols; ++jj) {
            int64_t j = int64_t(b) * bcols + jj;
            for (int t = 0; t < nnz_per_col; ++t) {
                trow.push_back(int64_t(b) * brows + int64_t(srng.below(uint64_t(brows))));
                tcol.push_back(j);
                tval.push_back(vrng.normal());
            }
            if (link_rows > 0 && srng.uniform() < link_fraction)
                for (int t = 0; t < link_per_col; ++t) {
                    trow.push_back(link_base + int64_t(srng.below(uint64_t(link_rows))));
                    tcol.push_back(j);
                    tval.push_back(vrng.normal());
                }
        }
    build_csr(p, trow, tcol, tval);
    std::vector<double> s;
    plant_columns(p, upper, zero_fraction, xrng, s);
    std::vector<int8_t> kind(p.rows, kEqual);
    std::vector<double> y0(p.rows);
    for (int64_t i = 0; i < p.rows; ++i) {
        if (i >= link_base) kind[i] = yrng.uniform() < 0.5 ? kTight : kSlack;
        double u = yrng.uniform(-1.0, 1.0);
        y0[i] = kind[i] == kEqual ? u : kind[i] == kTight ? -std::abs(u) : 0.0;
    }
    plant(p, kind, y0, s, yrng);
    p.name = "blockangular_" + std::to_string(blocks) + "x" + std::to_string(brows) + "x" +
             std::to_string(bcols) + "_s" + std::to_string(seed);
    return p;
}

// Shapes of the batched workload: m log-uniform in [m_lo, m_hi], n uniform in
// [2m, 3m], 3..8 nonzeros per column, box [0, U] with U ~ U(1, 50).
struct BatchShape {
    int64_t m, n;
    double upper;
    uint64_t seed;
};
inline std::vector<BatchShape> batch_shapes(int count, uint64_t seed, int64_t m_lo = 50,
                                            int64_t m_hi = 2000) {
    Rng rng(splitmix(seed ^ 0x71));
    std::vector<BatchShape> out(count);
    for (int i = 0; i < count; ++i) {
        double lm = rng.uniform(std::log(double(m_lo)), std::log(double(m_hi)));
        int64_t m = std::max<int64_t>(m_lo, std::min<int64_t>(m_hi, int64_t(std::llround(std::exp(lm)))));
        int64_t n = 2 * m + int64_t(rng.below(uint64_t(m + 1)));
        double upper = rng.uniform(1.0, 50.0);
        out[i] = {m, n, upper, splitmix(seed * 0x100000001b3ULL + uint64_t(i) + 1)};
    }
    return out;
}
inline Instance batch_instance(const BatchShape &s) {
    auto p = equality_lp(s.m, s.n, 3, 8, s.upper, s.seed);
    p.name = "batch_" + std::to_string(s.m) + "x" + std::to_string(s.n) + "_s" + std::to_string(s.seed);
    return p;
}

// Serving stream: request `index` draws its size log-uniformly in nonzeros
// over [1e3, 5e6] (about 13.75 nonzeros per row with n in [2m, 3m] and 3-8
// per column) and is otherwise a batch instance.
inline BatchShape serve_shape(int64_t index, uint64_t seed) {
    Rng rng(splitmix(seed ^ 0x73 ^ (uint64_t(index) * 0x9E3779B97F4A7C15ULL)));
    const double target = std::exp(rng.uniform(std::log(1e3), std::log(5e6)));
    const int64_t m = std::max<int64_t>(10, int64_t(std::llround(target / 13.75)));
    const int64_t n = 2 * m + int64_t(rng.below(uint64_t(m + 1)));
    const double upper = rng.uniform(1.0, 50.0);
    return {m, n, upper, splitmix(seed * 0x100000001b3ULL + uint64_t(index) + 0x5e7)};
}

// Named workloads.  `scale` divides the row/column counts (1 = full size) so
// tests can run the same structure small.
inline Instance workload(int config, uint64_t seed, int64_t scale = 1) {
    if (scale < 1) throw std::invalid_argument("scale must be >= 1");
    switch (config) {
    case 0:
        return equality_lp(std::max<int64_t>(2, 2000 / scale), std::max<int64_t>(4, 5000 / scale), 5, 5,
                           10, seed);
    case 1:
        return powerlaw_lp(std::max<int64_t>(10, 1000000 / scale), std::max<int64_t>(20, 2000000 / scale),
                           seed);
    case 4: {
        int blocks = int(std::max<int64_t>(2, 64 / std::min<int64_t>(scale, 32)));
        int64_t brows = std::max<int64_t>(10, 20000 / scale), bcols = std::max<int64_t>(20, 40000 / scale);
        return block_angular_lp(blocks, brows, bcols, 6, std::max<int64_t>(4, 2000 / scale), 0.10, 57,
                                seed);
    }
    default:
        throw std::invalid_argument("workload: config must be 0, 1 or 4 (2 and 3 are batches)");
    }
}

}  // namespace nyk_gen
