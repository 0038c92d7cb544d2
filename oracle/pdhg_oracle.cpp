// CPU oracle for the restarted-PDHG LP path.
//
// TEST INFRASTRUCTURE ONLY: this is the checker, a formula-for-formula CPU
// restatement of the reference's FP64 path.  Only the tests, the smoke check
// and the benchmark's CPU-baseline leg load it; the product never does.
//
// Every function follows the reference in operation order, OpenMP thresholds
// (`if (> 10000)`) and long-double guards, so that at one thread it reproduces
// the reference bit for bit.  Scope: continuous LPs (no Q, no integers), the
// pdhg / halpern / rhpdhg / r2hpdhg methods, Ruiz and combined scaling, the
// displacement and PID primal-weight rules.  Not restated: checkpoint files,
// feasibility polishing, power-iteration norms and GPU presolve (none is on by
// default in the reference).
#include "pdhg_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {
constexpr double inf = std::numeric_limits<double>::infinity();
using Clock = std::chrono::steady_clock;
static double elapsed(Clock::time_point s) { return std::chrono::duration<double>(Clock::now() - s).count(); }

struct Entry {
    int64_t row, col;
    double value;
};

// include/vantage/vantage.hpp:17-24
struct Sparse {
    int64_t rows = 0, cols = 0;
    std::vector<int64_t> ptr, index;
    std::vector<double> value;

    // src/model.cpp:11-38: sort COO by (row, col), sum duplicates, drop zeros.
    static Sparse build(int64_t m, int64_t n, std::vector<Entry> e) {
        if (m < 0 || n < 0) throw std::runtime_error("Negative matrix dimension");
        for (auto a : e)
            if (a.row < 0 || a.row >= m || a.col < 0 || a.col >= n || !std::isfinite(a.value))
                throw std::runtime_error("Invalid sparse entry");
        std::stable_sort(e.begin(), e.end(),
                         [](const Entry &a, const Entry &b) { return a.row < b.row || (a.row == b.row && a.col < b.col); });
        Sparse s;
        s.rows = m;
        s.cols = n;
        s.ptr.assign(m + 1, 0);
        for (size_t k = 0; k < e.size();) {
            auto a = e[k++];
            while (k < e.size() && e[k].row == a.row && e[k].col == a.col) a.value += e[k++].value;
            if (!std::isfinite(a.value)) throw std::runtime_error("Coefficient sum overflow");
            if (a.value != 0) {
                s.ptr[a.row + 1]++;
                s.index.push_back(a.col);
                s.value.push_back(a.value);
            }
        }
        for (int64_t i = 0; i < m; i++) s.ptr[i + 1] += s.ptr[i];
        return s;
    }
    // src/model.cpp:39-58
    Sparse transpose() const {
        Sparse t;
        t.rows = cols;
        t.cols = rows;
        t.ptr.assign(cols + 1, 0);
        t.index.resize(value.size());
        t.value.resize(value.size());
        for (auto j : index) t.ptr[j + 1]++;
        for (int64_t j = 0; j < cols; j++) t.ptr[j + 1] += t.ptr[j];
        auto pos = t.ptr;
        for (int64_t i = 0; i < rows; i++)
            for (auto k = ptr[i]; k < ptr[i + 1]; k++) {
                auto p = pos[index[k]]++;
                t.index[p] = i;
                t.value[p] = value[k];
            }
        return t;
    }
    // src/model.cpp:59-73
    std::vector<double> multiply(const std::vector<double> &x) const {
        if (x.size() != size_t(cols)) throw std::runtime_error("SpMV dimension mismatch");
        std::vector<double> y(rows);
#ifdef _OPENMP
#pragma omp parallel for if (rows > 10000)
#endif
        for (int64_t i = 0; i < rows; i++) {
            double v = 0;
            for (auto k = ptr[i]; k < ptr[i + 1]; k++) v += value[k] * x[index[k]];
            y[i] = v;
        }
        return y;
    }
};

// include/vantage/vantage.hpp:26-39, LP subset (q kept as an all-zero vector so
// that expressions mentioning it evaluate exactly as in the reference).
struct Model {
    Sparse A;
    std::vector<double> c, q, lb, ub, rl, ru;
    double offset = 0;
    // src/model.cpp:74-109 (LP checks; names are not carried by the oracle)
    void validate() const {
        auto n = c.size(), m = rl.size();
        if (q.size() != n || lb.size() != n || ub.size() != n || ru.size() != m || A.rows != int64_t(m) ||
            A.cols != int64_t(n))
            throw std::runtime_error("Inconsistent model dimensions");
        if (A.ptr.size() != m + 1 || A.ptr.front() != 0 || A.ptr.back() != int64_t(A.value.size()) ||
            A.index.size() != A.value.size())
            throw std::runtime_error("Invalid CSR storage");
        if (!std::isfinite(offset)) throw std::runtime_error("Invalid objective offset/sense");
        for (size_t j = 0; j < n; j++) {
            if (!std::isfinite(c[j]) || !std::isfinite(q[j])) throw std::runtime_error("Nonfinite objective coefficient");
            if (std::isnan(lb[j]) || std::isnan(ub[j]) || lb[j] == inf || ub[j] == -inf || lb[j] > ub[j])
                throw std::runtime_error("Invalid bounds for column " + std::to_string(j));
        }
        for (size_t i = 0; i < m; i++) {
            if (std::isnan(rl[i]) || std::isnan(ru[i]) || rl[i] == inf || ru[i] == -inf || rl[i] > ru[i])
                throw std::runtime_error("Invalid row bounds");
            if (A.ptr[i] > A.ptr[i + 1] || A.ptr[i] < 0 || A.ptr[i + 1] > int64_t(A.value.size()))
                throw std::runtime_error("Invalid CSR offsets");
        }
        for (size_t k = 0; k < A.value.size(); k++)
            if (A.index[k] < 0 || A.index[k] >= int64_t(n) || !std::isfinite(A.value[k]))
                throw std::runtime_error("Invalid matrix coefficient/index");
    }
};

struct Options {
    double tol = 1e-6, time_limit = 60;
    int64_t iteration_limit = 100000;
    int check_every = 100, scaling_passes = 5, threads = 1;
    bool presolve = true, restart = true, adaptive = true;
    std::string method = "pdhg", scaling = "ruiz", primal_weight = "displacement";
    std::vector<double> initial_x, initial_y;
};

struct Accuracy {
    double objective = inf, primal = inf, dual = inf, gap = inf, kkt = inf, integrality = inf;
    double primal_absolute = inf, dual_absolute = inf, complementarity = inf, lower_bound = -inf;
    bool finite = false;
};

// ---------------------------------------------------------------- verifier --
// src/verify.cpp:6-11
static long double down(long double v) { return std::nextafter(v, -std::numeric_limits<long double>::infinity()); }
static long double up(long double v) { return std::nextafter(v, std::numeric_limits<long double>::infinity()); }

// src/verify.cpp:12-88 (LP: no curvature point)
static double lp_lower_bound(const Model &m, const Sparse &at, const std::vector<double> &y, bool feasibility) {
    long double lower = feasibility ? 0 : m.offset;
    for (size_t i = 0; i < y.size(); i++)
        if (y[i] != 0) {
            double b = y[i] > 0 ? m.ru[i] : m.rl[i];
            if (!std::isfinite(b)) return -inf;
            if (b == 0) continue;
            long double term = (long double)y[i] * b;
            lower = down(lower - up(term));
        }
    for (size_t j = 0; j < m.c.size(); j++) {
        long double gl = feasibility ? 0 : m.c[j], gu = gl;
        for (auto k = at.ptr[j]; k < at.ptr[j + 1]; k++)
            if (y[at.index[k]] != 0 && at.value[k] != 0) {
                long double v = (long double)at.value[k] * y[at.index[k]];
                gl = down(gl + down(v));
                gu = up(gu + up(v));
            }
        if ((!std::isfinite(m.lb[j]) && gu > 0) || (!std::isfinite(m.ub[j]) && gl < 0)) return -inf;
        long double term = std::numeric_limits<long double>::infinity();
        if (gl <= 0 && gu >= 0) term = 0;
        for (double b : {m.lb[j], m.ub[j]})
            if (std::isfinite(b)) {
                term = std::min(term, (b == 0 || gl == 0) ? 0 : down(gl * b));
                term = std::min(term, (b == 0 || gu == 0) ? 0 : down(gu * b));
            }
        if (gl == 0 && gu == 0) term = 0;
        if (!std::isfinite(term)) return -inf;
        if (term != 0) lower = down(lower + term);
    }
    double result = double(lower);
    if (!std::isfinite(result)) return -inf;
    return std::nextafter(result, -inf);
}

// src/verify.cpp:90-240
class Verifier {
    const Model &model;
    Sparse transpose;
    std::vector<double> activity, transpose_product;

  public:
    explicit Verifier(const Model &source)
        : model(source), transpose(source.A.transpose()), activity(source.A.rows), transpose_product(source.A.cols) {}

    double infeasibility_bound(const std::vector<double> &y) const {
        if (y.size() != model.rl.size()) return -inf;
        for (auto v : y)
            if (!std::isfinite(v)) return -inf;
        return lp_lower_bound(model, transpose, y, true);
    }

    Accuracy evaluate(const std::vector<double> &x, const std::vector<double> &y, bool compute_safe_bound = true) {
        const auto &m = model;
        Accuracy a;
        if (x.size() != m.c.size() || y.size() != m.rl.size()) return a;
        for (auto v : x)
            if (!std::isfinite(v)) return a;
        for (auto v : y)
            if (!std::isfinite(v)) return a;
        auto multiply_into = [](const Sparse &s, const std::vector<double> &input, std::vector<double> &output) {
#ifdef _OPENMP
#pragma omp parallel for if (s.rows > 10000)
#endif
            for (int64_t i = 0; i < s.rows; i++) {
                double sum = 0;
                for (auto k = s.ptr[i]; k < s.ptr[i + 1]; k++) sum += s.value[k] * input[s.index[k]];
                output[i] = sum;
            }
        };
        multiply_into(m.A, x, activity);
        multiply_into(transpose, y, transpose_product);
        const auto &ax = activity;
        const auto &aty = transpose_product;
        for (auto v : ax)
            if (!std::isfinite(v)) return Accuracy{};
        for (auto v : aty)
            if (!std::isfinite(v)) return Accuracy{};
        a.objective = m.offset;
        a.primal = 0;
        a.dual = 0;
        a.integrality = 0;
        a.primal_absolute = 0;
        a.dual_absolute = 0;
        long double support = 0, box_min = 0, complement = 0;
        bool bound_ok = true;
        for (size_t i = 0; i < y.size(); i++) {
            double v = std::max({0., m.rl[i] - ax[i], ax[i] - m.ru[i]});
            if (std::isfinite(m.rl[i])) a.primal = std::max(a.primal, std::max(0., m.rl[i] - ax[i]) / (1 + std::abs(m.rl[i])));
            if (std::isfinite(m.ru[i])) a.primal = std::max(a.primal, std::max(0., ax[i] - m.ru[i]) / (1 + std::abs(m.ru[i])));
            a.primal_absolute = std::max(a.primal_absolute, v);
            double endpoint = y[i] > 0 ? m.ru[i] : m.rl[i];
            if (y[i] != 0) {
                if (std::isfinite(endpoint)) {
                    support += (long double)y[i] * endpoint;
                    complement += std::abs((long double)y[i] * (endpoint - ax[i]));
                } else {
                    bound_ok = false;
                    a.dual_absolute = std::max(a.dual_absolute, std::abs(y[i]));
                    a.dual = std::max(a.dual, std::abs(y[i]));
                }
            }
        }
        const double qx = 0;  // LP: Q x == 0
        double cscale = 1;
        for (auto v : m.c) cscale = std::max(cscale, 1 + std::abs(v));
        for (size_t j = 0; j < x.size(); j++) {
            double v = std::max({0., m.lb[j] - x[j], x[j] - m.ub[j]});
            double bs = 1 + std::abs(x[j]);
            a.primal = std::max(a.primal, v / bs);
            a.primal_absolute = std::max(a.primal_absolute, v);
            a.objective += m.c[j] * x[j] + .5 * m.q[j] * x[j] * x[j] + .5 * x[j] * qx;
            double linear = m.c[j] + aty[j], g = linear + m.q[j] * x[j] + qx;
            if (!std::isfinite(linear) || !std::isfinite(g)) return Accuracy{};
            double d = std::clamp(g, x[j] - m.ub[j], x[j] - m.lb[j]);
            a.dual_absolute = std::max(a.dual_absolute, std::abs(d));
            a.dual = std::max(a.dual, std::abs(d) / cscale);
            if (g > 0 && std::isfinite(m.lb[j])) complement += std::abs((long double)g * (x[j] - m.lb[j]));
            if (g < 0 && std::isfinite(m.ub[j])) complement += std::abs((long double)g * (m.ub[j] - x[j]));
            double arg = 0;
            if (m.q[j] > 0)
                arg = std::clamp(-linear / m.q[j], m.lb[j], m.ub[j]);
            else if (linear > 0)
                arg = m.lb[j];
            else if (linear < 0)
                arg = m.ub[j];
            else
                arg = std::clamp(0., m.lb[j], m.ub[j]);
            if (!std::isfinite(arg))
                bound_ok = false;
            else
                box_min += (long double)linear * arg + .5L * m.q[j] * arg * arg;
        }
        a.complementarity = double(complement);
        a.gap = a.complementarity / (1 + std::abs(a.objective - m.offset));
        if (bound_ok) {
            long double lower = (long double)m.offset - support + box_min;
            long double guard = 64 * std::numeric_limits<double>::epsilon() *
                                (1 + std::abs((long double)m.offset) + std::abs(support) + std::abs(box_min));
            a.lower_bound = double(lower - guard);
            a.gap = std::max(a.gap, std::abs(a.objective - a.lower_bound) / (1 + std::abs(a.objective) + std::abs(a.lower_bound)));
        }
        if (compute_safe_bound)
            a.lower_bound = lp_lower_bound(m, transpose, y, false);
        else
            a.lower_bound = -inf;
        if (std::isfinite(a.lower_bound))
            a.gap = std::max(a.gap, std::abs(a.objective - a.lower_bound) / (1 + std::abs(a.objective) + std::abs(a.lower_bound)));
        a.kkt = std::max({a.primal, a.dual, a.gap});
        a.finite = std::isfinite(a.objective) && std::isfinite(a.kkt) && std::isfinite(a.integrality) &&
                   std::isfinite(a.gap) && std::isfinite(a.primal) && std::isfinite(a.dual);
        if (!a.finite) {
            a.kkt = inf;
            a.lower_bound = -inf;
        }
        return a;
    }
};

// ---------------------------------------------------------------- prepare --
// src/internal.hpp:16-25
struct Prepared {
    Model model;
    std::vector<int64_t> cols, rows, lower_rows, upper_rows;
    std::vector<double> fixed, column_scale, row_scale;
    double objective_scale = 1;
    std::string failure, reason;
    std::vector<double> failure_ray;
    // src/preprocess.cpp:121-126
    std::vector<double> restore_x(const std::vector<double> &x) const {
        auto out = fixed;
        for (size_t j = 0; j < cols.size(); j++) out[cols[j]] = x[j] * column_scale[j];
        return out;
    }
    // src/preprocess.cpp:127-132
    std::vector<double> restore_y(const std::vector<double> &y, size_t n) const {
        std::vector<double> out(n);
        for (size_t i = 0; i < rows.size(); i++)
            out[y[i] >= 0 ? upper_rows[i] : lower_rows[i]] += y[i] * row_scale[i] * objective_scale;
        return out;
    }
};

// src/preprocess.cpp:6-9
static bool representable(double before, double after) {
    return std::isfinite(before) ? std::isfinite(after) && (before == 0 || after != 0) : before == after;
}
// src/preprocess.cpp:12-65 (LP: Q empty)
static void extra_scale(Prepared &p, std::vector<double> row, std::vector<double> col) {
    auto &m = p.model;
    for (size_t i = 0; i < row.size(); i++) {
        double f = row[i];
        bool safe = representable(p.row_scale[i], p.row_scale[i] * f) && representable(m.rl[i], m.rl[i] * f) &&
                    representable(m.ru[i], m.ru[i] * f);
        for (auto k = m.A.ptr[i]; k < m.A.ptr[i + 1]; k++) safe &= representable(m.A.value[k], m.A.value[k] * f);
        if (!safe) continue;
        p.row_scale[i] *= f;
        m.rl[i] *= f;
        m.ru[i] *= f;
        for (auto k = m.A.ptr[i]; k < m.A.ptr[i + 1]; k++) m.A.value[k] *= f;
    }
    for (size_t j = 0; j < col.size(); j++) {
        double f = col[j];
        if (!representable(p.column_scale[j], p.column_scale[j] * f) || !representable(m.c[j], m.c[j] * f) ||
            !representable(m.q[j], m.q[j] * (f * f)) || !representable(m.lb[j], m.lb[j] / f) ||
            !representable(m.ub[j], m.ub[j] / f))
            col[j] = 1;
    }
    for (size_t k = 0; k < m.A.value.size(); k++) {
        auto j = m.A.index[k];
        if (!representable(m.A.value[k], m.A.value[k] * col[j])) col[j] = 1;
    }
    for (size_t j = 0; j < col.size(); j++) {
        double f = col[j];
        p.column_scale[j] *= f;
        m.c[j] *= f;
        m.q[j] *= f * f;
        m.lb[j] /= f;
        m.ub[j] /= f;
    }
    for (size_t k = 0; k < m.A.value.size(); k++) m.A.value[k] *= col[m.A.index[k]];
}
// src/preprocess.cpp:66-72
static double geometric_factor(double low, double high) {
    if (high == 0) return 1;
    return std::exp(std::clamp(-.5 * (std::log(low) + std::log(high)), std::log(1e-3), std::log(1e3)));
}
// src/preprocess.cpp:73-101
static void geometric_scale(Prepared &p) {
    auto &m = p.model;
    std::vector<double> row(m.rl.size(), 1), col(m.c.size(), 1);
    for (size_t i = 0; i < row.size(); i++) {
        double low = inf, high = 0;
        for (auto k = m.A.ptr[i]; k < m.A.ptr[i + 1]; k++) {
            double v = std::abs(m.A.value[k]);
            if (v > 0) {
                low = std::min(low, v);
                high = std::max(high, v);
            }
        }
        row[i] = geometric_factor(low, high);
    }
    extra_scale(p, row, col);
    std::vector<double> low(col.size(), inf), high(col.size());
    for (size_t k = 0; k < m.A.value.size(); k++) {
        auto j = m.A.index[k];
        double v = std::abs(m.A.value[k]);
        if (v > 0) {
            low[j] = std::min(low[j], v);
            high[j] = std::max(high[j], v);
        }
    }
    for (size_t j = 0; j < col.size(); j++) col[j] = geometric_factor(low[j], high[j]);
    std::fill(row.begin(), row.end(), 1);
    extra_scale(p, row, col);
}
// src/preprocess.cpp:102-119
static void pock_chambolle_scale(Prepared &p) {
    auto &m = p.model;
    std::vector<long double> rows(m.rl.size()), cols(m.c.size());
    for (size_t i = 0; i < rows.size(); i++)
        for (auto k = m.A.ptr[i]; k < m.A.ptr[i + 1]; k++) {
            long double v = std::abs(m.A.value[k]);
            rows[i] += v;
            cols[m.A.index[k]] += v;
        }
    auto factor = [](long double sum) {
        return sum > 0 && std::isfinite(sum) ? double(std::clamp(1 / std::sqrt(sum), 1e-3L, 1e3L)) : 1.;
    };
    std::vector<double> row(rows.size()), col(cols.size());
    std::transform(rows.begin(), rows.end(), row.begin(), factor);
    std::transform(cols.begin(), cols.end(), col.begin(), factor);
    extra_scale(p, row, col);
}

// src/preprocess.cpp:133-346 (LP path; no GPU compaction, no time/interrupt stops)
static Prepared prepare(const Model &src, const Options &o) {
    Prepared p;
    p.fixed.assign(src.c.size(), 0);
    Model &m = p.model;
    m.offset = src.offset;
    std::vector<int64_t> map(src.c.size(), -1);
    std::vector<bool> used(src.c.size(), false);
    for (auto j : src.A.index) used[j] = true;
    for (size_t j = 0; j < src.c.size(); j++) {
        double chosen = src.lb[j];
        bool removable = src.lb[j] == src.ub[j];
        if (!used[j]) {
            chosen = src.q[j] > 0   ? std::clamp(-src.c[j] / src.q[j], src.lb[j], src.ub[j])
                     : src.c[j] > 0 ? src.lb[j]
                     : src.c[j] < 0 ? src.ub[j]
                                    : std::clamp(0., src.lb[j], src.ub[j]);
            removable = std::isfinite(chosen);
        }
        if (o.presolve && removable) {
            p.fixed[j] = chosen;
            m.offset += src.c[j] * chosen + .5 * src.q[j] * chosen * chosen;
        } else {
            map[j] = int64_t(p.cols.size());
            p.cols.push_back(int64_t(j));
            m.c.push_back(src.c[j]);
            m.q.push_back(src.q[j]);
            m.lb.push_back(src.lb[j]);
            m.ub.push_back(src.ub[j]);
        }
    }
    std::vector<Entry> e;
    std::map<std::vector<std::pair<int64_t, double>>, size_t> duplicate_rows;
    for (size_t i = 0; i < src.rl.size(); i++) {
        long double shift = 0, shift_magnitude = 0;
        int64_t count = 0;
        for (auto k = src.A.ptr[i]; k < src.A.ptr[i + 1]; k++) {
            auto j = src.A.index[k];
            if (map[j] < 0) {
                long double term = (long double)src.A.value[k] * p.fixed[j];
                shift += term;
                shift_magnitude += std::abs(term);
            } else
                count++;
        }
        double l = src.rl[i] - double(shift), u = src.ru[i] - double(shift);
        if (count == 0 && o.presolve) {
            long double guard = 8 * std::numeric_limits<long double>::epsilon() * (src.A.ptr[i + 1] - src.A.ptr[i] + 1) *
                                (1 + shift_magnitude);
            if (shift + guard < src.rl[i] || shift - guard > src.ru[i]) {
                p.failure = "INFEASIBLE";
                p.reason = "Empty row after fixed-variable substitution: r" + std::to_string(i);
                return p;
            }
            if (l <= 0 && u >= 0) continue;
        }
        long double amin = 0, amax = 0, magnitude = 0;
        for (auto k = src.A.ptr[i]; k < src.A.ptr[i + 1]; k++) {
            auto j = src.A.index[k];
            double v = src.A.value[k];
            amin += (long double)v * (v > 0 ? src.lb[j] : src.ub[j]);
            amax += (long double)v * (v > 0 ? src.ub[j] : src.lb[j]);
            magnitude += std::abs((long double)v * (v > 0 ? src.lb[j] : src.ub[j])) +
                         std::abs((long double)v * (v > 0 ? src.ub[j] : src.lb[j]));
        }
        if (o.presolve) {
            long double slack = 8 * std::numeric_limits<long double>::epsilon() * (src.A.ptr[i + 1] - src.A.ptr[i] + 1) *
                                (1 + magnitude);
            if (std::isfinite(amin) && std::isfinite(amax) && (amin > src.ru[i] + slack || amax < src.rl[i] - slack)) {
                p.failure = "INFEASIBLE";
                p.reason = "Verified disjoint row activity bounds: r" + std::to_string(i);
                return p;
            }
        }
        if (o.presolve && std::isfinite(amin) && std::isfinite(amax)) {
            long double guard = 8 * std::numeric_limits<long double>::epsilon() * (src.A.ptr[i + 1] - src.A.ptr[i] + 1) *
                                (1 + magnitude);
            if (amin - guard >= src.rl[i] && amax + guard <= src.ru[i]) continue;
        }
        if (o.presolve && count == src.A.ptr[i + 1] - src.A.ptr[i]) {
            std::vector<std::pair<int64_t, double>> key;
            for (auto k = src.A.ptr[i]; k < src.A.ptr[i + 1]; ++k) key.push_back({map[src.A.index[k]], src.A.value[k]});
            auto found = duplicate_rows.find(key);
            if (found != duplicate_rows.end()) {
                auto r = found->second;
                if (l > m.rl[r]) {
                    m.rl[r] = l;
                    p.lower_rows[r] = int64_t(i);
                }
                if (u < m.ru[r]) {
                    m.ru[r] = u;
                    p.upper_rows[r] = int64_t(i);
                }
                if (m.rl[r] > m.ru[r]) {
                    p.failure = "INFEASIBLE";
                    p.reason = "Disjoint bounds on identical constraint rows";
                    p.failure_ray.assign(src.rl.size(), 0);
                    p.failure_ray[p.lower_rows[r]] = -1;
                    p.failure_ray[p.upper_rows[r]] += 1;
                    if (!(Verifier(src).infeasibility_bound(p.failure_ray) > 0)) p.failure_ray.clear();
                    return p;
                }
                continue;
            }
            duplicate_rows.emplace(std::move(key), p.rows.size());
        }
        auto row = p.rows.size();
        p.rows.push_back(int64_t(i));
        p.lower_rows.push_back(int64_t(i));
        p.upper_rows.push_back(int64_t(i));
        m.rl.push_back(l);
        m.ru.push_back(u);
        for (auto k = src.A.ptr[i]; k < src.A.ptr[i + 1]; k++)
            if (map[src.A.index[k]] >= 0) e.push_back({int64_t(row), map[src.A.index[k]], src.A.value[k]});
    }
    m.A = Sparse::build(int64_t(m.rl.size()), int64_t(m.c.size()), std::move(e));
    p.column_scale.assign(m.c.size(), 1);
    p.row_scale.assign(m.rl.size(), 1);
    if (o.scaling == "combined" && o.scaling_passes > 0) geometric_scale(p);
    for (int pass = 0; pass < o.scaling_passes; pass++) {
        std::vector<double> row(m.rl.size(), 1), col(m.c.size(), 1);
        for (size_t i = 0; i < row.size(); ++i) {
            double norm = 0;
            for (auto k = m.A.ptr[i]; k < m.A.ptr[i + 1]; ++k) norm = std::max(norm, std::abs(m.A.value[k]));
            row[i] = norm > 0 ? std::clamp(1 / std::sqrt(norm), 1e-3, 1e3) : 1.;
        }
        extra_scale(p, row, col);
        std::vector<double> norm(col.size());
        for (size_t k = 0; k < m.A.value.size(); ++k)
            norm[m.A.index[k]] = std::max(norm[m.A.index[k]], std::abs(m.A.value[k]));
        for (size_t j = 0; j < col.size(); ++j) col[j] = norm[j] > 0 ? std::clamp(1 / std::sqrt(norm[j]), 1e-3, 1e3) : 1.;
        std::fill(row.begin(), row.end(), 1);
        extra_scale(p, row, col);
    }
    if (o.scaling == "combined" && o.scaling_passes > 0) pock_chambolle_scale(p);
    if (o.scaling_passes > 0) {
        double cnorm = 0;
        for (auto c : m.c) cnorm = std::max(cnorm, std::abs(c));
        for (auto q : m.q) cnorm = std::max(cnorm, std::abs(q));
        p.objective_scale = cnorm > 0 ? cnorm : 1;
        for (auto &c : m.c) c /= p.objective_scale;
        for (auto &q : m.q) q /= p.objective_scale;
        m.offset /= p.objective_scale;
    }
    m.validate();
    return p;
}

// ---------------------------------------------------------------- backend --
// src/cpu.cpp:4-248 (LP: no Q safety scaling)
class CpuBackend {
    const Model &m;
    Sparse at;
    std::vector<double> x, y, delta_x, xa, ya, gradient, trial_x, trial_y, ax, next_ax;
    bool adaptive, halpern;
    std::vector<double> anchor_x, anchor_y, anchor_ax;
    int64_t epoch_steps = 0;
    double reflection;
    bool fixed_point_restarts, request_restart = false;
    double initial_residual = inf, previous_residual = inf;
    double factor = 1, average_mass = 0;
    int64_t accepted = 0, rejected = 0;

  public:
    CpuBackend(const Model &model, const std::vector<double> &px, const std::vector<double> &py, bool adapt,
               bool anchored, double reflect, bool residual_restarts)
        : m(model), at(m.A.transpose()), x(px), y(py), delta_x(px), xa(px.size()), ya(py.size()),
          gradient(px.size()), trial_x(px.size()), trial_y(py.size()), ax(m.A.multiply(px)), next_ax(py.size()),
          adaptive(adapt), halpern(anchored), reflection(reflect), fixed_point_restarts(residual_restarts) {
        if (halpern) {
            anchor_x = x;
            anchor_y = y;
            anchor_ax = ax;
        }
    }
    int64_t rejected_steps() const { return rejected; }
    bool restart_requested() const { return request_restart; }
    // src/cpu.cpp:96-222
    int advance(int count, double primal_step, double dual_step) {
        for (int it = 0; it < count; it++) {
#ifdef _OPENMP
#pragma omp parallel for if (m.A.cols > 10000)
#endif
            for (int64_t j = 0; j < m.A.cols; j++) {
                double g = m.c[j];
                for (auto k = at.ptr[j]; k < at.ptr[j + 1]; k++) g += at.value[k] * y[at.index[k]];
                gradient[j] = g;
            }
            bool success = false;
            for (int attempt = 0; attempt < 40; attempt++) {
                double tau = primal_step * factor, sigma = dual_step * factor;
                double dx2 = 0, dy2 = 0, coupling = 0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : dx2) if (m.A.cols > 10000)
#endif
                for (int64_t j = 0; j < m.A.cols; j++) {
                    trial_x[j] = std::clamp((x[j] - tau * gradient[j]) / (1 + tau * m.q[j]), m.lb[j], m.ub[j]);
                    double dx = trial_x[j] - x[j];
                    dx2 += dx * dx;
                    delta_x[j] = dx;
                }
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : dy2, coupling) if (m.A.rows > 10000)
#endif
                for (int64_t i = 0; i < m.A.rows; i++) {
                    double delta = 0;
                    for (auto k = m.A.ptr[i]; k < m.A.ptr[i + 1]; k++) delta += m.A.value[k] * delta_x[m.A.index[k]];
                    double bar = ax[i] + 2 * delta;
                    double z = y[i] + sigma * bar;
                    trial_y[i] = std::max(0., z - sigma * m.ru[i]) + std::min(0., z - sigma * m.rl[i]);
                    double dy = trial_y[i] - y[i];
                    dy2 += dy * dy;
                    if (!std::isfinite(bar) || !std::isfinite(trial_y[i])) dy2 = inf;
                    coupling += dy * delta;
                    next_ax[i] = ax[i] + delta;
                }
                double energy = dx2 / tau + dy2 / sigma;
                double limit = std::abs(coupling) > 0 ? energy / (2 * std::abs(coupling)) : inf;
                bool finite = std::isfinite(energy) && std::isfinite(coupling);
                double used_factor = factor;
                bool accept = finite && (!adaptive || limit >= 1);
                if (adaptive) {
                    double k = double(accepted + 2);
                    double multiplier = std::min((1 - std::pow(k, -.3)) * limit, 1 + std::pow(k, -.6));
                    if (!finite || !std::isfinite(multiplier) || multiplier <= 0) multiplier = .5;
                    factor = std::clamp(factor * multiplier, 1e-12, 1e12);
                }
                if (!accept) {
                    rejected++;
                    if (!adaptive) throw std::runtime_error("Nonfinite PDHG step");
                    continue;
                }
                x.swap(trial_x);
                y.swap(trial_y);
                ax.swap(next_ax);
                accepted++;
                if (halpern) {
                    xa = x;
                    ya = y;
                    double lambda = 1 / (double(epoch_steps) + 2);
                    for (size_t j = 0; j < x.size(); j++)
                        x[j] = lambda * anchor_x[j] + (1 - lambda) * ((1 + reflection) * x[j] - reflection * trial_x[j]);
                    for (size_t i = 0; i < y.size(); i++) {
                        y[i] = lambda * anchor_y[i] + (1 - lambda) * ((1 + reflection) * y[i] - reflection * trial_y[i]);
                        ax[i] = lambda * anchor_ax[i] + (1 - lambda) * ((1 + reflection) * ax[i] - reflection * next_ax[i]);
                    }
                    double residual = std::sqrt(std::max(0., energy - 2 * coupling));
                    if (epoch_steps == 0) initial_residual = residual;
                    epoch_steps++;
                    request_restart = fixed_point_restarts && epoch_steps >= 10 &&
                                      (residual <= .2 * initial_residual ||
                                       (residual <= .8 * initial_residual && residual > previous_residual) ||
                                       epoch_steps >= std::max<int64_t>(10, accepted / 2));
                    previous_residual = residual;
                    if (request_restart) return it + 1;
                    success = true;
                    break;
                }
                average_mass += used_factor;
                double w = used_factor / average_mass;
#ifdef _OPENMP
#pragma omp parallel for if (m.A.cols > 10000)
#endif
                for (int64_t j = 0; j < m.A.cols; j++) xa[j] += (x[j] - xa[j]) * w;
#ifdef _OPENMP
#pragma omp parallel for if (m.A.rows > 10000)
#endif
                for (int64_t i = 0; i < m.A.rows; i++) ya[i] += (y[i] - ya[i]) * w;
                success = true;
                break;
            }
            if (!success) throw std::runtime_error("PDHG backtracking exhausted");
        }
        return count;
    }
    // src/cpu.cpp:223-231
    void candidates(std::vector<double> &cx, std::vector<double> &cy, std::vector<double> &avgx,
                    std::vector<double> &avgy) {
        cx = x;
        cy = y;
        avgx = xa;
        avgy = ya;
        ax = m.A.multiply(x);
    }
    // src/cpu.cpp:232-247
    void reset(const std::vector<double> &px, const std::vector<double> &py) {
        x = px;
        y = py;
        ax = m.A.multiply(x);
        if (halpern) {
            anchor_x = x;
            anchor_y = y;
            anchor_ax = ax;
            epoch_steps = 0;
            initial_residual = previous_residual = inf;
            request_restart = false;
        }
        average_mass = 0;
        std::fill(xa.begin(), xa.end(), 0);
        std::fill(ya.begin(), ya.end(), 0);
    }
};

// src/research.cpp:52-70
class PrimalWeightController {
    double integral = 0, previous = 0;
    bool initialized = false;

  public:
    double update(double weight, long double dx, long double dy) {
        if (!(weight > 0) || !std::isfinite(weight) || !(dx > 1e-24L) || !(dy > 1e-24L) || !std::isfinite(dx) ||
            !std::isfinite(dy))
            return weight;
        double error = double(std::log((long double)weight) + .5L * (std::log(dx) - std::log(dy)));
        error = std::clamp(error, -10., 10.);
        double next_integral = std::clamp(integral + error, -100., 100.);
        double derivative = initialized ? error - previous : 0;
        double correction = std::clamp(.3 * error + .01 * next_integral + .1 * derivative, -std::log(4.), std::log(4.));
        double logarithm = std::log(weight) - correction;
        double clipped = std::clamp(logarithm, std::log(1e-4), std::log(1e4));
        if (clipped == logarithm) integral = next_integral;
        previous = error;
        initialized = true;
        return std::exp(clipped);
    }
};

struct Result {
    std::string status = "UNKNOWN", message;
    std::vector<double> x, y, infeasibility_ray;
    Accuracy accuracy;
    double seconds = 0, preprocess_seconds = 0, iteration_seconds = 0, verification_seconds = 0;
    int64_t iterations = 0, restarts = 0, rejected_steps = 0, removed_rows = 0, removed_columns = 0,
            weight_updates = 0, host_candidate_checks = 0;
    double certificate_margin = 0;
    // decision trace (not in the reference's Result): first step and weight,
    // first restart iteration and the weight right after it
    double initial_step = 0, initial_weight = 0, first_restart_weight = 0;
    int64_t first_restart_iteration = 0;
};

static bool anchored_method(const Options &o) { return o.method == "halpern" || o.method == "rhpdhg" || o.method == "r2hpdhg"; }

// src/solver.cpp:18-54 (LP PDHG subset)
static void validate_options(const Options &o) {
    if (!(o.tol > 0 && o.tol < 1) || !std::isfinite(o.tol) || o.iteration_limit < 0 || o.time_limit < 0 ||
        std::isnan(o.time_limit) || o.check_every < 1 || o.scaling_passes < 0 || o.scaling_passes > 20 || o.threads < 1)
        throw std::runtime_error("Invalid solver options");
    if (o.method != "pdhg" && !anchored_method(o))
        throw std::runtime_error("Unknown method: use pdhg, halpern, rhpdhg or r2hpdhg");
    if (o.primal_weight != "displacement" && o.primal_weight != "pid")
        throw std::runtime_error("Unknown primal weight controller");
    if (o.scaling != "ruiz" && o.scaling != "combined") throw std::runtime_error("Unknown scaling: use ruiz or combined");
}

// src/solver.cpp:106-735, pdhg branch (lines 215-735) on the CPU backend.
static Result solve_continuous(const Model &original, const Options &o) {
    if (anchored_method(o) && o.adaptive)
        throw std::runtime_error("Experimental Halpern supports continuous LP with --no-adaptive");
    auto start = Clock::now();
    Result r;
    Verifier verifier(original);
    r.status = "ITERATION_LIMIT";
    auto prep = prepare(original, o);
    r.preprocess_seconds = elapsed(start);
    if (!prep.failure.empty()) {
        r.status = prep.failure;
        r.message = prep.reason;
        r.infeasibility_ray = prep.failure_ray;
        if (!r.infeasibility_ray.empty()) r.certificate_margin = Verifier(original).infeasibility_bound(r.infeasibility_ray);
        r.seconds = elapsed(start);
        return r;
    }
    auto &m = prep.model;
    r.removed_columns = int64_t(original.c.size() - m.c.size());
    r.removed_rows = int64_t(original.rl.size() - m.rl.size());
    std::vector<double> x(m.c.size()), y(m.rl.size());
    for (size_t j = 0; j < x.size(); j++) {
        auto col = prep.cols[j];
        double initial = o.initial_x.empty() ? 0 : o.initial_x[col];
        x[j] = std::clamp(initial / prep.column_scale[j], m.lb[j], m.ub[j]);
    }
    for (size_t i = 0; i < y.size(); i++)
        if (!o.initial_y.empty()) y[i] = o.initial_y[prep.rows[i]] / (prep.row_scale[i] * prep.objective_scale);
    r.x = prep.restore_x(x);
    r.y = prep.restore_y(y, original.rl.size());
    r.accuracy = verifier.evaluate(r.x, r.y);
    if (r.accuracy.finite && r.accuracy.primal_absolute == 0) {
        std::vector<bool> used(m.c.size(), false);
        for (auto j : m.A.index) used[j] = true;
        for (size_t j = 0; j < x.size(); j++)
            if (!used[j] && m.q[j] == 0 && ((m.c[j] < 0 && m.ub[j] == inf) || (m.c[j] > 0 && m.lb[j] == -inf))) {
                r.status = "UNBOUNDED";
                r.message = "Feasible point and isolated improving recession direction";
                r.seconds = elapsed(start);
                return r;
            }
    }
    if (r.accuracy.finite && r.accuracy.kkt <= o.tol) {
        r.status = "OPTIMAL";
        r.seconds = elapsed(start);
        return r;
    }
    if (o.time_limit == 0) {
        r.status = "TIME_LIMIT";
        r.seconds = elapsed(start);
        return r;
    }
    std::vector<double> colsum(x.size());
    double rowmax = 0;
    for (size_t i = 0; i < y.size(); i++) {
        double s = 0;
        for (auto k = m.A.ptr[i]; k < m.A.ptr[i + 1]; k++) {
            double v = std::abs(m.A.value[k]);
            s += v;
            colsum[m.A.index[k]] += v;
        }
        rowmax = std::max(rowmax, s);
    }
    double colmax = 0;
    for (auto s : colsum) colmax = std::max(colmax, s);
    double norm = std::sqrt(rowmax) * std::sqrt(colmax);
    double step = norm > 0 ? .9 / norm : 1, weight = 1;
    if ((o.adaptive || anchored_method(o)) && o.primal_weight != "pid") {
        long double c2 = 0, b2 = 0;
        for (auto c : m.c) c2 += (long double)c * c;
        for (size_t i = 0; i < m.rl.size(); i++) {
            double b = std::clamp(0., m.rl[i], m.ru[i]);
            b2 += (long double)b * b;
        }
        if (c2 > 1e-24 && b2 > 1e-24) weight = std::clamp(std::sqrt(double(c2 / b2)), 1e-4, 1e4);
    }
    r.initial_step = step;
    r.initial_weight = weight;
    bool residual_restarts = o.method == "rhpdhg" || o.method == "r2hpdhg";
    CpuBackend backend(m, x, y, o.adaptive, anchored_method(o), o.method == "r2hpdhg" ? 1 : 0,
                       residual_restarts && o.restart);
    double restart_kkt = r.accuracy.kkt;
    int64_t since_restart = 0;
    std::vector<double> epochx = x, epochy = y;
    PrimalWeightController controller;
    int64_t previous_rejected = 0;
    while (r.iterations < o.iteration_limit) {
        if (elapsed(start) >= o.time_limit) {
            r.status = "TIME_LIMIT";
            break;
        }
        int count = int(std::min<int64_t>(o.check_every, o.iteration_limit - r.iterations));
        auto tick = Clock::now();
        int accepted = backend.advance(count, step / weight, step * weight);
        r.iteration_seconds += elapsed(tick);
        r.iterations += accepted;
        since_restart += accepted;
        auto rejected = backend.rejected_steps();
        r.rejected_steps += rejected - previous_rejected;
        previous_rejected = rejected;
        r.host_candidate_checks++;
        tick = Clock::now();
        std::vector<double> cx, cy, ax, ay;
        backend.candidates(cx, cy, ax, ay);
        std::vector<double> restart_x, restart_y;
        if (residual_restarts && backend.restart_requested()) {
            restart_x = ax;
            restart_y = ay;
        }
        auto ox = prep.restore_x(cx);
        auto oy = prep.restore_y(cy, original.rl.size());
        auto ca = verifier.evaluate(ox, oy, false);
        auto aox = prep.restore_x(ax);
        auto aoy = prep.restore_y(ay, original.rl.size());
        auto aa = verifier.evaluate(aox, aoy, false);
        bool avg = aa.kkt < ca.kkt;
        if (avg) {
            cx.swap(ax);
            cy.swap(ay);
            ox.swap(aox);
            oy.swap(aoy);
            ca = aa;
        }
        if (ca.finite && ca.kkt < r.accuracy.kkt) {
            r.x = ox;
            r.y = oy;
            r.accuracy = ca;
        }
        r.verification_seconds += elapsed(tick);
        if (!ca.finite) {
            r.status = "NUMERICAL_ERROR";
            r.message = "Nonfinite iterate or diagnostic";
            break;
        }
        if (ca.kkt <= o.tol) ca = verifier.evaluate(ox, oy);
        if (ca.kkt <= o.tol) {
            r.x = std::move(ox);
            r.y = std::move(oy);
            r.accuracy = ca;
            r.status = "OPTIMAL";
            break;
        }
        if (r.iterations >= 1000 && ca.primal > o.tol) {
            auto displacement = cy;
            for (size_t i = 0; i < displacement.size(); ++i) displacement[i] -= epochy[i];
            std::vector<std::vector<double>> rays;
            rays.push_back(oy);
            rays.push_back(aoy);
            rays.push_back(prep.restore_y(displacement, original.rl.size()));
            for (auto &ray : rays) {
                double norm_y = 0;
                bool finite = true;
                for (double v : ray) {
                    finite = finite && std::isfinite(v);
                    norm_y = std::max(norm_y, std::abs(v));
                }
                if (!finite || !(norm_y > 0)) continue;
                for (double &v : ray) v /= norm_y;
                double margin = verifier.infeasibility_bound(ray);
                if (margin > 0 && std::isfinite(margin)) {
                    r.status = "INFEASIBLE";
                    r.message = "Independently verified box/row Farkas certificate";
                    r.infeasibility_ray = std::move(ray);
                    r.certificate_margin = margin;
                    break;
                }
            }
            if (r.status == "INFEASIBLE") break;
        }
        bool restart_due = residual_restarts ? backend.restart_requested()
                                             : since_restart >= int64_t(o.check_every) * 2 &&
                                                   (ca.kkt < .5 * restart_kkt ||
                                                    since_restart >= std::max<int64_t>(2000, r.iterations / 2));
        if (o.restart && restart_due) {
            if (residual_restarts) {
                cx = std::move(restart_x);
                cy = std::move(restart_y);
            }
            if (o.adaptive || anchored_method(o) || o.primal_weight == "pid") {
                long double dx = 0, dy = 0;
                for (size_t j = 0; j < cx.size(); j++) dx += (long double)(cx[j] - epochx[j]) * (cx[j] - epochx[j]);
                for (size_t i = 0; i < cy.size(); i++) dy += (long double)(cy[i] - epochy[i]) * (cy[i] - epochy[i]);
                if (dx > 1e-24 && dy > 1e-24) {
                    if (o.primal_weight == "pid")
                        weight = controller.update(weight, dx, dy);
                    else {
                        double target = std::clamp(std::sqrt(double(dy / dx)), 1e-4, 1e4);
                        weight = std::sqrt(weight * target);
                    }
                    r.weight_updates++;
                }
            }
            backend.reset(cx, cy);
            epochx = cx;
            epochy = cy;
            restart_kkt = ca.kkt;
            since_restart = 0;
            if (r.restarts++ == 0) {
                r.first_restart_iteration = r.iterations;
                r.first_restart_weight = weight;
            }
        }
    }
    auto check = Clock::now();
    r.accuracy = verifier.evaluate(r.x, r.y);
    r.verification_seconds += elapsed(check);
    if (r.status == "OPTIMAL" && (!r.accuracy.finite || r.accuracy.kkt > o.tol)) r.status = "NUMERICAL_ERROR";
    r.seconds = elapsed(start);
    return r;
}

// src/solver.cpp:55-105 (LP PDHG subset)
static Result solve(const Model &m, const Options &o) {
    auto overall_start = Clock::now();
    m.validate();
    validate_options(o);
    if (anchored_method(o) && o.adaptive)
        throw std::runtime_error("Experimental Halpern supports continuous LP with --no-adaptive");
    if (!o.initial_x.empty() && o.initial_x.size() != m.c.size())
        throw std::runtime_error("Warm-start primal dimension mismatch");
    if (!o.initial_y.empty() && o.initial_y.size() != m.rl.size())
        throw std::runtime_error("Warm-start dual dimension mismatch");
    for (auto v : o.initial_x)
        if (!std::isfinite(v)) throw std::runtime_error("Nonfinite warm start");
    for (auto v : o.initial_y)
        if (!std::isfinite(v)) throw std::runtime_error("Nonfinite warm start");
#ifdef _OPENMP
    omp_set_num_threads(o.threads);
#endif
    auto result = solve_continuous(m, o);
    result.seconds = elapsed(overall_start);
    return result;
}

static Model to_model(const oracle_lp *lp) {
    Model m;
    m.A.rows = lp->rows;
    m.A.cols = lp->cols;
    int64_t nnz = lp->ptr[lp->rows];
    m.A.ptr.assign(lp->ptr, lp->ptr + lp->rows + 1);
    m.A.index.assign(lp->index, lp->index + nnz);
    m.A.value.assign(lp->value, lp->value + nnz);
    m.c.assign(lp->c, lp->c + lp->cols);
    m.q.assign(lp->cols, 0);
    m.lb.assign(lp->lb, lp->lb + lp->cols);
    m.ub.assign(lp->ub, lp->ub + lp->cols);
    m.rl.assign(lp->rl, lp->rl + lp->rows);
    m.ru.assign(lp->ru, lp->ru + lp->rows);
    m.offset = lp->offset;
    return m;
}
static Options to_options(const oracle_lp *lp, const oracle_options *in) {
    Options o;
    if (!in) return o;
    o.tol = in->tol;
    o.time_limit = in->time_limit;
    o.iteration_limit = in->iteration_limit;
    o.check_every = in->check_every;
    o.scaling_passes = in->scaling_passes;
    o.threads = in->threads;
    o.presolve = in->presolve != 0;
    o.restart = in->restart != 0;
    o.adaptive = in->adaptive != 0;
    if (in->method) o.method = in->method;
    if (in->scaling) o.scaling = in->scaling;
    if (in->primal_weight) o.primal_weight = in->primal_weight;
    if (in->initial_x) o.initial_x.assign(in->initial_x, in->initial_x + lp->cols);
    if (in->initial_y) o.initial_y.assign(in->initial_y, in->initial_y + lp->rows);
    return o;
}
static void copy_accuracy(const Accuracy &a, oracle_accuracy *o) {
    o->objective = a.objective;
    o->primal = a.primal;
    o->dual = a.dual;
    o->gap = a.gap;
    o->kkt = a.kkt;
    o->lower_bound = a.lower_bound;
    o->complementarity = a.complementarity;
    o->primal_absolute = a.primal_absolute;
    o->dual_absolute = a.dual_absolute;
    o->finite = a.finite ? 1 : 0;
}
struct BackendHandle {
    Model model;
    std::unique_ptr<CpuBackend> backend;
};
static void set_error(char *err, int len, const char *msg) {
    if (err && len > 0) std::snprintf(err, size_t(len), "%s", msg);
}
}  // namespace oracle

extern "C" {
void oracle_default_options(oracle_options *o) {
    o->tol = 1e-6;
    o->time_limit = 60;
    o->iteration_limit = 100000;
    o->check_every = 100;
    o->scaling_passes = 5;
    o->threads = 1;
    o->presolve = 1;
    o->restart = 1;
    o->adaptive = 1;
    o->method = "pdhg";
    o->scaling = "ruiz";
    o->primal_weight = "displacement";
    o->initial_x = nullptr;
    o->initial_y = nullptr;
}

int oracle_solve(const oracle_lp *lp, const oracle_options *opts, oracle_result *out, char *err, int errlen) {
    try {
        auto m = oracle::to_model(lp);
        auto o = oracle::to_options(lp, opts);
        auto r = oracle::solve(m, o);
        std::snprintf(out->status, sizeof(out->status), "%s", r.status.c_str());
        oracle::copy_accuracy(r.accuracy, &out->accuracy);
        out->iterations = r.iterations;
        out->restarts = r.restarts;
        out->rejected_steps = r.rejected_steps;
        out->weight_updates = r.weight_updates;
        out->host_candidate_checks = r.host_candidate_checks;
        out->removed_rows = r.removed_rows;
        out->removed_columns = r.removed_columns;
        out->seconds = r.seconds;
        out->iteration_seconds = r.iteration_seconds;
        out->preprocess_seconds = r.preprocess_seconds;
        out->verification_seconds = r.verification_seconds;
        out->initial_step = r.initial_step;
        out->initial_weight = r.initial_weight;
        out->first_restart_weight = r.first_restart_weight;
        out->first_restart_iteration = r.first_restart_iteration;
        if (out->x && r.x.size() == size_t(lp->cols)) std::copy(r.x.begin(), r.x.end(), out->x);
        if (out->y && r.y.size() == size_t(lp->rows)) std::copy(r.y.begin(), r.y.end(), out->y);
        return 0;
    } catch (const std::exception &e) {
        oracle::set_error(err, errlen, e.what());
        return -1;
    }
}

int oracle_verify(const oracle_lp *lp, const double *x, const double *y, int safe, oracle_accuracy *a, char *err,
                  int errlen) {
    try {
        auto m = oracle::to_model(lp);
        oracle::Verifier v(m);
        auto acc = v.evaluate(std::vector<double>(x, x + lp->cols), std::vector<double>(y, y + lp->rows), safe != 0);
        oracle::copy_accuracy(acc, a);
        return 0;
    } catch (const std::exception &e) {
        oracle::set_error(err, errlen, e.what());
        return -1;
    }
}

double oracle_infeasibility_bound(const oracle_lp *lp, const double *ray) {
    auto m = oracle::to_model(lp);
    oracle::Verifier v(m);
    return v.infeasibility_bound(std::vector<double>(ray, ray + lp->rows));
}

int oracle_prepare(const oracle_lp *lp, const oracle_options *opts, int64_t *out_rows, int64_t *out_cols,
                   double *column_scale, double *row_scale, double *objective_scale, int64_t *col_map,
                   int64_t *row_map, char *failure, int failure_len) {
    try {
        auto m = oracle::to_model(lp);
        auto o = oracle::to_options(lp, opts);
        auto p = oracle::prepare(m, o);
        oracle::set_error(failure, failure_len, p.failure.c_str());
        if (!p.failure.empty()) return 1;
        *out_rows = int64_t(p.rows.size());
        *out_cols = int64_t(p.cols.size());
        std::copy(p.column_scale.begin(), p.column_scale.end(), column_scale);
        std::copy(p.row_scale.begin(), p.row_scale.end(), row_scale);
        *objective_scale = p.objective_scale;
        std::copy(p.cols.begin(), p.cols.end(), col_map);
        std::copy(p.rows.begin(), p.rows.end(), row_map);
        return 0;
    } catch (const std::exception &e) {
        oracle::set_error(failure, failure_len, e.what());
        return -1;
    }
}

int oracle_prepared_model(const oracle_lp *lp, const oracle_options *opts, int64_t *out_rows, int64_t *out_cols,
                          int64_t *ptr, int64_t *index, double *value, double *c, double *lb, double *ub,
                          double *rl, double *ru, double *offset, char *failure, int failure_len) {
    try {
        auto m = oracle::to_model(lp);
        auto o = oracle::to_options(lp, opts);
        auto p = oracle::prepare(m, o);
        oracle::set_error(failure, failure_len, p.failure.c_str());
        if (!p.failure.empty()) return 1;
        const auto &q = p.model;
        *out_rows = q.A.rows;
        *out_cols = q.A.cols;
        std::copy(q.A.ptr.begin(), q.A.ptr.end(), ptr);
        std::copy(q.A.index.begin(), q.A.index.end(), index);
        std::copy(q.A.value.begin(), q.A.value.end(), value);
        std::copy(q.c.begin(), q.c.end(), c);
        std::copy(q.lb.begin(), q.lb.end(), lb);
        std::copy(q.ub.begin(), q.ub.end(), ub);
        std::copy(q.rl.begin(), q.rl.end(), rl);
        std::copy(q.ru.begin(), q.ru.end(), ru);
        *offset = q.offset;
        return 0;
    } catch (const std::exception &e) {
        oracle::set_error(failure, failure_len, e.what());
        return -1;
    }
}

void *oracle_backend_create(const oracle_lp *lp, const double *x0, const double *y0, int adaptive, int halpern,
                            double reflection, int residual_restarts, char *err, int errlen) {
    try {
        if (halpern && adaptive) throw std::runtime_error("Halpern requires a fixed PDHG operator");
        if (!std::isfinite(reflection) || reflection < 0 || reflection > 1)
            throw std::runtime_error("Invalid Halpern reflection");
        auto h = new oracle::BackendHandle;
        h->model = oracle::to_model(lp);
        h->backend = std::make_unique<oracle::CpuBackend>(
            h->model, std::vector<double>(x0, x0 + lp->cols), std::vector<double>(y0, y0 + lp->rows), adaptive != 0,
            halpern != 0, reflection, residual_restarts != 0);
        return h;
    } catch (const std::exception &e) {
        oracle::set_error(err, errlen, e.what());
        return nullptr;
    }
}
int oracle_backend_advance(void *b, int count, double tau, double sigma, char *err, int errlen) {
    try {
        return static_cast<oracle::BackendHandle *>(b)->backend->advance(count, tau, sigma);
    } catch (const std::exception &e) {
        oracle::set_error(err, errlen, e.what());
        return -1;
    }
}
void oracle_backend_candidates(void *b, double *x, double *y, double *xa, double *ya) {
    auto h = static_cast<oracle::BackendHandle *>(b);
    std::vector<double> cx, cy, ax, ay;
    h->backend->candidates(cx, cy, ax, ay);
    std::copy(cx.begin(), cx.end(), x);
    std::copy(cy.begin(), cy.end(), y);
    std::copy(ax.begin(), ax.end(), xa);
    std::copy(ay.begin(), ay.end(), ya);
}
void oracle_backend_reset(void *b, const double *x, const double *y) {
    auto h = static_cast<oracle::BackendHandle *>(b);
    h->backend->reset(std::vector<double>(x, x + h->model.A.cols), std::vector<double>(y, y + h->model.A.rows));
}
int64_t oracle_backend_rejected(void *b) { return static_cast<oracle::BackendHandle *>(b)->backend->rejected_steps(); }
int oracle_backend_restart_requested(void *b) {
    return static_cast<oracle::BackendHandle *>(b)->backend->restart_requested() ? 1 : 0;
}
void oracle_backend_destroy(void *b) { delete static_cast<oracle::BackendHandle *>(b); }

void *oracle_pid_create(void) { return new oracle::PrimalWeightController(); }
double oracle_pid_update(void *pid, double weight, double dx, double dy) {
    return static_cast<oracle::PrimalWeightController *>(pid)->update(weight, (long double)dx, (long double)dy);
}
void oracle_pid_destroy(void *pid) { delete static_cast<oracle::PrimalWeightController *>(pid); }
}
