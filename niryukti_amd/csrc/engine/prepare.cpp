// Presolve and scaling of the LP before it is uploaded to the GPU.
//
// The arithmetic follows the reference's prepare() (src/preprocess.cpp:133-346)
// exactly: the same presolve rules, the same long-double guards and the same
// clipped Ruiz / geometric / Pock-Chambolle factors, so the scaled LP handed to
// the device is bit-identical to the reference's.  The implementation differs
// where it matters at 1M-row scale: duplicate rows are found with a hash table
// instead of an ordered map of whole rows, and the per-row and per-column
// factor passes run in parallel (every factor is a max or a per-index product,
// so the result does not depend on the thread count).
#include "lp.hpp"
#include "parallel.hpp"

#include <algorithm>
#include <stdexcept>
#include <unordered_map>

namespace nyk {

std::vector<double> Prepared::restore_x(const std::vector<double> &x) const {
    auto out = fixed;
    for (size_t j = 0; j < cols.size(); ++j) out[cols[j]] = x[j] * column_scale[j];
    return out;
}

std::vector<double> Prepared::restore_y(const std::vector<double> &y, size_t n) const {
    std::vector<double> out(n);
    for (size_t i = 0; i < rows.size(); ++i)
        out[y[i] >= 0 ? upper_rows[i] : lower_rows[i]] += y[i] * row_scale[i] * objective_scale;
    return out;
}

namespace {

// A scaling step may never erase a nonzero or turn a finite value infinite
// (src/preprocess.cpp:6-9).
inline bool keeps(double before, double after) {
    return std::isfinite(before) ? std::isfinite(after) && (before == 0 || after != 0) : before == after;
}

// Applies row factors then column factors, skipping any row or column whose
// transformation is unsafe as a whole (src/preprocess.cpp:12-65).
void apply_scaling(Prepared &p, const std::vector<double> &row, std::vector<double> col) {
    Model &m = p.model;
    const int64_t rows = m.rows(), cols = m.cols();
    const bool any_row = std::any_of(row.begin(), row.end(), [](double f) { return f != 1; });
    if (any_row)
        parallel_for(rows, rows > 4096, [&](int64_t i) {
            double f = row[i];
            if (f == 1) return;  // multiplying by one changes nothing
            bool safe = keeps(p.row_scale[i], p.row_scale[i] * f) && keeps(m.rl[i], m.rl[i] * f) &&
                        keeps(m.ru[i], m.ru[i] * f);
            for (int64_t k = m.A.ptr[i]; k < m.A.ptr[i + 1] && safe; ++k) safe = keeps(m.A.value[k], m.A.value[k] * f);
            if (!safe) return;
            p.row_scale[i] *= f;
            m.rl[i] *= f;
            m.ru[i] *= f;
            for (int64_t k = m.A.ptr[i]; k < m.A.ptr[i + 1]; ++k) m.A.value[k] *= f;
        });
    if (!std::any_of(col.begin(), col.end(), [](double f) { return f != 1; })) return;
    for (int64_t j = 0; j < cols; ++j) {
        double f = col[j];
        if (f == 1) continue;
        if (!keeps(p.column_scale[j], p.column_scale[j] * f) || !keeps(m.c[j], m.c[j] * f) ||
            !keeps(0.0, 0.0 * (f * f)) || !keeps(m.lb[j], m.lb[j] / f) || !keeps(m.ub[j], m.ub[j] / f))
            col[j] = 1;
    }
    // A column whose factor would erase or overflow any coefficient keeps factor 1.
    std::vector<uint8_t> unsafe(cols, 0);
    for (int64_t k = 0; k < m.A.nnz(); ++k) {
        int64_t j = m.A.index[k];
        if (!keeps(m.A.value[k], m.A.value[k] * col[j])) unsafe[j] = 1;
    }
    for (int64_t j = 0; j < cols; ++j) {
        if (unsafe[j]) col[j] = 1;
        double f = col[j];
        p.column_scale[j] *= f;
        m.c[j] *= f;
        m.lb[j] /= f;
        m.ub[j] /= f;
    }
    parallel_for(m.A.nnz(), m.A.nnz() > 65536, [&](int64_t k) { m.A.value[k] *= col[m.A.index[k]]; });
}

// src/preprocess.cpp:66-72
double geometric_factor(double low, double high) {
    if (high == 0) return 1;
    return std::exp(std::clamp(-.5 * (std::log(low) + std::log(high)), std::log(1e-3), std::log(1e3)));
}

// src/preprocess.cpp:73-101
void geometric_pass(Prepared &p) {
    Model &m = p.model;
    std::vector<double> row(m.rows(), 1), ones_c(m.cols(), 1);
    parallel_for(m.rows(), m.rows() > 4096, [&](int64_t i) {
        double low = kInf, high = 0;
        for (int64_t k = m.A.ptr[i]; k < m.A.ptr[i + 1]; ++k) {
            double v = std::abs(m.A.value[k]);
            if (v > 0) {
                low = std::min(low, v);
                high = std::max(high, v);
            }
        }
        row[i] = geometric_factor(low, high);
    });
    apply_scaling(p, row, ones_c);
    std::vector<double> low(m.cols(), kInf), high(m.cols(), 0);
    for (int64_t k = 0; k < m.A.nnz(); ++k) {
        int64_t j = m.A.index[k];
        double v = std::abs(m.A.value[k]);
        if (v > 0) {
            low[j] = std::min(low[j], v);
            high[j] = std::max(high[j], v);
        }
    }
    std::vector<double> col(m.cols());
    for (int64_t j = 0; j < m.cols(); ++j) col[j] = geometric_factor(low[j], high[j]);
    apply_scaling(p, std::vector<double>(m.rows(), 1), col);
}

// src/preprocess.cpp:102-119: long-double absolute row and column sums,
// accumulated in storage order.
void pock_chambolle_pass(Prepared &p) {
    Model &m = p.model;
    std::vector<long double> rs(m.rows(), 0), cs(m.cols(), 0);
    for (int64_t i = 0; i < m.rows(); ++i)
        for (int64_t k = m.A.ptr[i]; k < m.A.ptr[i + 1]; ++k) {
            long double v = std::abs(m.A.value[k]);
            rs[i] += v;
            cs[m.A.index[k]] += v;
        }
    auto factor = [](long double sum) {
        return sum > 0 && std::isfinite(sum) ? double(std::clamp(1 / std::sqrt(sum), 1e-3L, 1e3L)) : 1.;
    };
    std::vector<double> row(m.rows()), col(m.cols());
    std::transform(rs.begin(), rs.end(), row.begin(), factor);
    std::transform(cs.begin(), cs.end(), col.begin(), factor);
    apply_scaling(p, row, col);
}

// Ruiz infinity-norm equilibration pass (src/preprocess.cpp:306-324).
void ruiz_pass(Prepared &p) {
    Model &m = p.model;
    std::vector<double> row(m.rows(), 1);
    parallel_for(m.rows(), m.rows() > 4096, [&](int64_t i) {
        double norm = 0;
        for (int64_t k = m.A.ptr[i]; k < m.A.ptr[i + 1]; ++k) norm = std::max(norm, std::abs(m.A.value[k]));
        row[i] = norm > 0 ? std::clamp(1 / std::sqrt(norm), 1e-3, 1e3) : 1.;
    });
    apply_scaling(p, row, std::vector<double>(m.cols(), 1));
    std::vector<double> norm(m.cols(), 0);
    for (int64_t k = 0; k < m.A.nnz(); ++k) norm[m.A.index[k]] = std::max(norm[m.A.index[k]], std::abs(m.A.value[k]));
    std::vector<double> col(m.cols());
    for (int64_t j = 0; j < m.cols(); ++j) col[j] = norm[j] > 0 ? std::clamp(1 / std::sqrt(norm[j]), 1e-3, 1e3) : 1.;
    apply_scaling(p, std::vector<double>(m.rows(), 1), col);
}

struct RowKeyHash {
    size_t operator()(const std::vector<std::pair<int64_t, double>> &key) const {
        uint64_t h = 1469598103934665603ULL;
        for (const auto &e : key) {
            h = (h ^ uint64_t(e.first)) * 1099511628211ULL;
            uint64_t bits;
            static_assert(sizeof bits == sizeof e.second);
            __builtin_memcpy(&bits, &e.second, sizeof bits);
            h = (h ^ bits) * 1099511628211ULL;
        }
        return size_t(h);
    }
};

}  // namespace

// src/preprocess.cpp:133-346 (continuous LP path).
Prepared prepare(const Model &src, const Options &o) {
    Prepared p;
    const int64_t n0 = src.cols(), m0 = src.rows();
    p.fixed.assign(n0, 0);
    Model &m = p.model;
    m.name = src.name;
    m.offset = src.offset;
    m.sense = src.sense;
    std::vector<int64_t> map(n0, -1);
    std::vector<uint8_t> used(n0, 0);
    for (auto j : src.A.index) used[j] = 1;
    // Fixed columns and columns that appear in no row leave the LP
    // (src/preprocess.cpp:159-184).
    for (int64_t j = 0; j < n0; ++j) {
        double chosen = src.lb[j];
        bool removable = src.lb[j] == src.ub[j];
        if (!used[j]) {
            chosen = src.c[j] > 0 ? src.lb[j] : src.c[j] < 0 ? src.ub[j] : std::clamp(0., src.lb[j], src.ub[j]);
            removable = std::isfinite(chosen);
        }
        if (o.presolve && removable) {
            p.fixed[j] = chosen;
            m.offset += src.c[j] * chosen + .5 * 0.0 * chosen * chosen;
        } else {
            map[j] = int64_t(p.cols.size());
            p.cols.push_back(j);
            m.c.push_back(src.c[j]);
            m.lb.push_back(src.lb[j]);
            m.ub.push_back(src.ub[j]);
            if (!src.names.empty()) m.names.push_back(src.names[j]);
        }
    }
    std::vector<Entry> entries;
    entries.reserve(src.A.value.size());
    std::unordered_map<std::vector<std::pair<int64_t, double>>, size_t, RowKeyHash> duplicate_rows;
    constexpr long double eps = std::numeric_limits<long double>::epsilon();
    for (int64_t i = 0; i < m0; ++i) {
        const int64_t begin = src.A.ptr[i], end = src.A.ptr[i + 1], len = end - begin;
        long double shift = 0, shift_magnitude = 0;
        int64_t count = 0;
        for (int64_t k = begin; k < end; ++k) {
            int64_t j = src.A.index[k];
            if (map[j] < 0) {
                long double term = (long double)src.A.value[k] * p.fixed[j];
                shift += term;
                shift_magnitude += std::abs(term);
            } else
                ++count;
        }
        double l = src.rl[i] - double(shift), u = src.ru[i] - double(shift);
        // Rows left empty by substitution: infeasible or dropped (:216-226).
        if (count == 0 && o.presolve) {
            long double guard = 8 * eps * (len + 1) * (1 + shift_magnitude);
            if (shift + guard < src.rl[i] || shift - guard > src.ru[i]) {
                p.failure = "INFEASIBLE";
                p.reason = "Empty row after fixed-variable substitution: " + src.row_name(i);
                return p;
            }
            if (l <= 0 && u >= 0) continue;
        }
        // Activity bounds over the box prove infeasibility or redundancy (:227-252).
        long double amin = 0, amax = 0, magnitude = 0;
        for (int64_t k = begin; k < end; ++k) {
            int64_t j = src.A.index[k];
            double v = src.A.value[k];
            amin += (long double)v * (v > 0 ? src.lb[j] : src.ub[j]);
            amax += (long double)v * (v > 0 ? src.ub[j] : src.lb[j]);
            magnitude += std::abs((long double)v * (v > 0 ? src.lb[j] : src.ub[j])) +
                         std::abs((long double)v * (v > 0 ? src.ub[j] : src.lb[j]));
        }
        if (o.presolve) {
            long double slack = 8 * eps * (len + 1) * (1 + magnitude);
            if (std::isfinite(amin) && std::isfinite(amax) && (amin > src.ru[i] + slack || amax < src.rl[i] - slack)) {
                p.failure = "INFEASIBLE";
                p.reason = "Verified disjoint row activity bounds: " + src.row_name(i);
                return p;
            }
        }
        if (o.presolve && std::isfinite(amin) && std::isfinite(amax)) {
            long double guard = 8 * eps * (len + 1) * (1 + magnitude);
            if (amin - guard >= src.rl[i] && amax + guard <= src.ru[i]) continue;
        }
        // Exactly repeated rows are intersected into one (:253-283).
        if (o.presolve && count == len) {
            std::vector<std::pair<int64_t, double>> key;
            key.reserve(size_t(len));
            for (int64_t k = begin; k < end; ++k) key.emplace_back(map[src.A.index[k]], src.A.value[k]);
            auto found = duplicate_rows.find(key);
            if (found != duplicate_rows.end()) {
                size_t r = found->second;
                if (l > m.rl[r]) {
                    m.rl[r] = l;
                    p.lower_rows[r] = i;
                }
                if (u < m.ru[r]) {
                    m.ru[r] = u;
                    p.upper_rows[r] = i;
                }
                if (m.rl[r] > m.ru[r]) {
                    p.failure = "INFEASIBLE";
                    p.reason = "Disjoint bounds on identical constraint rows";
                    p.failure_ray.assign(m0, 0);
                    p.failure_ray[p.lower_rows[r]] = -1;
                    p.failure_ray[p.upper_rows[r]] += 1;
                    if (!(Verifier(src).infeasibility_bound(p.failure_ray) > 0)) p.failure_ray.clear();
                    return p;
                }
                continue;
            }
            duplicate_rows.emplace(std::move(key), p.rows.size());
        }
        const int64_t row = int64_t(p.rows.size());
        p.rows.push_back(i);
        p.lower_rows.push_back(i);
        p.upper_rows.push_back(i);
        m.rl.push_back(l);
        m.ru.push_back(u);
        if (!src.row_names.empty()) m.row_names.push_back(src.row_names[i]);
        for (int64_t k = begin; k < end; ++k)
            if (map[src.A.index[k]] >= 0) entries.push_back({row, map[src.A.index[k]], src.A.value[k]});
    }
    m.A = Csr::from_entries(int64_t(m.rl.size()), int64_t(m.c.size()), std::move(entries));
    p.column_scale.assign(m.c.size(), 1);
    p.row_scale.assign(m.rl.size(), 1);
    const bool combined = o.scaling == "combined" && o.scaling_passes > 0;
    if (combined) geometric_pass(p);
    for (int pass = 0; pass < o.scaling_passes; ++pass) ruiz_pass(p);
    if (combined) pock_chambolle_pass(p);
    if (o.scaling_passes > 0) {
        double cnorm = 0;
        for (double c : m.c) cnorm = std::max(cnorm, std::abs(c));
        p.objective_scale = cnorm > 0 ? cnorm : 1;
        for (double &c : m.c) c /= p.objective_scale;
        m.offset /= p.objective_scale;
    }
    m.validate();
    return p;
}

}  // namespace nyk
