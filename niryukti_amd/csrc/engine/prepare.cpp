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
#include <random>
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
    // Per-row analysis (src/preprocess.cpp:185-252) is independent from row
    // to row, so it runs in parallel; the decisions that depend on earlier
    // rows (the first failure, duplicate rows) are then taken in row order,
    // exactly as the sequential loop takes them.
    enum : uint8_t { kKeep, kDrop, kEmptyInfeasible, kDisjoint };
    struct RowInfo {
        double l, u;
        uint64_t hash;  // of (column, value) pairs, for rows without removed columns
        uint8_t action;
        bool keyed;     // presolve and count == len: candidate for duplicate detection
    };
    std::vector<RowInfo> info(static_cast<size_t>(m0));
    constexpr long double eps = std::numeric_limits<long double>::epsilon();
    parallel_for(m0, m0 > 4096, [&](int64_t i) {
        RowInfo &r = info[size_t(i)];
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
        r.l = src.rl[i] - double(shift);
        r.u = src.ru[i] - double(shift);
        r.action = kKeep;
        r.keyed = false;
        r.hash = 0;
        // Rows left empty by substitution: infeasible or dropped (:216-226).
        if (count == 0 && o.presolve) {
            long double guard = 8 * eps * (len + 1) * (1 + shift_magnitude);
            if (shift + guard < src.rl[i] || shift - guard > src.ru[i]) {
                r.action = kEmptyInfeasible;
                return;
            }
            if (r.l <= 0 && r.u >= 0) {
                r.action = kDrop;
                return;
            }
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
                r.action = kDisjoint;
                return;
            }
        }
        if (o.presolve && std::isfinite(amin) && std::isfinite(amax)) {
            long double guard = 8 * eps * (len + 1) * (1 + magnitude);
            if (amin - guard >= src.rl[i] && amax + guard <= src.ru[i]) {
                r.action = kDrop;
                return;
            }
        }
        if (o.presolve && count == len) {
            r.keyed = true;
            uint64_t h = 1469598103934665603ULL;
            for (int64_t k = begin; k < end; ++k) {
                h = (h ^ uint64_t(map[src.A.index[k]])) * 1099511628211ULL;
                uint64_t bits;
                __builtin_memcpy(&bits, &src.A.value[k], sizeof bits);
                h = (h ^ bits) * 1099511628211ULL;
            }
            r.hash = h;
        }
    });
    // Exactly repeated rows are intersected into one (:253-283): rows match
    // when their (column, value) sequences are equal.
    auto same_row = [&](int64_t a, int64_t b) {
        const int64_t la = src.A.ptr[a + 1] - src.A.ptr[a], lb = src.A.ptr[b + 1] - src.A.ptr[b];
        if (la != lb) return false;
        for (int64_t t = 0; t < la; ++t) {
            const int64_t ka = src.A.ptr[a] + t, kb = src.A.ptr[b] + t;
            if (map[src.A.index[ka]] != map[src.A.index[kb]] || !(src.A.value[ka] == src.A.value[kb])) return false;
        }
        return true;
    };
    // Only rows whose hash occurs more than once can repeat an earlier row.
    std::vector<uint8_t> shared_hash(static_cast<size_t>(m0), 0);
    {
        std::vector<std::pair<uint64_t, int64_t>> keyed;
        for (int64_t i = 0; i < m0; ++i)
            if (info[size_t(i)].action == kKeep && info[size_t(i)].keyed) keyed.emplace_back(info[size_t(i)].hash, i);
        std::sort(keyed.begin(), keyed.end());
        for (size_t a = 0; a < keyed.size();) {
            size_t b = a + 1;
            while (b < keyed.size() && keyed[b].first == keyed[a].first) ++b;
            if (b - a > 1)
                for (size_t t = a; t < b; ++t) shared_hash[size_t(keyed[t].second)] = 1;
            a = b;
        }
    }
    std::unordered_map<uint64_t, std::vector<std::pair<int64_t, size_t>>> duplicate_rows;  // hash -> (row, kept index)
    std::vector<int64_t> kept;
    kept.reserve(size_t(m0));
    for (int64_t i = 0; i < m0; ++i) {
        const RowInfo &r = info[size_t(i)];
        if (r.action == kEmptyInfeasible) {
            p.failure = "INFEASIBLE";
            p.reason = "Empty row after fixed-variable substitution: " + src.row_name(i);
            return p;
        }
        if (r.action == kDisjoint) {
            p.failure = "INFEASIBLE";
            p.reason = "Verified disjoint row activity bounds: " + src.row_name(i);
            return p;
        }
        if (r.action == kDrop) continue;
        const double l = r.l, u = r.u;
        if (r.keyed && shared_hash[size_t(i)]) {
            auto &bucket = duplicate_rows[r.hash];
            bool merged = false;
            for (const auto &[other, rr] : bucket) {
                if (!same_row(other, i)) continue;
                if (l > m.rl[rr]) {
                    m.rl[rr] = l;
                    p.lower_rows[rr] = i;
                }
                if (u < m.ru[rr]) {
                    m.ru[rr] = u;
                    p.upper_rows[rr] = i;
                }
                if (m.rl[rr] > m.ru[rr]) {
                    p.failure = "INFEASIBLE";
                    p.reason = "Disjoint bounds on identical constraint rows";
                    p.failure_ray.assign(m0, 0);
                    p.failure_ray[p.lower_rows[rr]] = -1;
                    p.failure_ray[p.upper_rows[rr]] += 1;
                    if (!(Verifier(src).infeasibility_bound(p.failure_ray) > 0)) p.failure_ray.clear();
                    return p;
                }
                merged = true;
                break;
            }
            if (merged) continue;
            bucket.emplace_back(i, p.rows.size());
        }
        p.rows.push_back(i);
        p.lower_rows.push_back(i);
        p.upper_rows.push_back(i);
        m.rl.push_back(l);
        m.ru.push_back(u);
        if (!src.row_names.empty()) m.row_names.push_back(src.row_names[i]);
        kept.push_back(i);
    }
    // The kept rows' entries, columns renumbered: already a canonical CSR when
    // every row's renumbered columns increase and no value is zero (the usual
    // case, since the source matrix is canonical); otherwise the general
    // builder sorts and merges (src/model.cpp:11-38).
    const int64_t rows_kept = int64_t(kept.size());
    bool canonical = true;
    int64_t total = 0;
    for (int64_t r = 0; r < rows_kept && canonical; ++r) {
        int64_t prev = -1;
        for (int64_t k = src.A.ptr[kept[r]]; k < src.A.ptr[kept[r] + 1]; ++k) {
            const int64_t j = map[src.A.index[k]];
            if (j < 0) continue;
            if (j <= prev || src.A.value[k] == 0 || !std::isfinite(src.A.value[k])) {
                canonical = false;
                break;
            }
            prev = j;
            ++total;
        }
    }
    if (canonical) {
        Csr &A = m.A;
        A.rows = rows_kept;
        A.cols = int64_t(m.c.size());
        A.ptr.assign(size_t(rows_kept) + 1, 0);
        for (int64_t r = 0; r < rows_kept; ++r) {
            int64_t c = 0;
            for (int64_t k = src.A.ptr[kept[r]]; k < src.A.ptr[kept[r] + 1]; ++k) c += map[src.A.index[k]] >= 0;
            A.ptr[size_t(r) + 1] = A.ptr[size_t(r)] + c;
        }
        A.index.resize(size_t(total));
        A.value.resize(size_t(total));
        parallel_for(rows_kept, rows_kept > 4096, [&](int64_t r) {
            int64_t at = A.ptr[size_t(r)];
            for (int64_t k = src.A.ptr[kept[r]]; k < src.A.ptr[kept[r] + 1]; ++k) {
                const int64_t j = map[src.A.index[k]];
                if (j < 0) continue;
                A.index[size_t(at)] = j;
                A.value[size_t(at)] = src.A.value[k];
                ++at;
            }
        });
    } else {
        std::vector<Entry> entries;
        entries.reserve(src.A.value.size());
        for (int64_t r = 0; r < rows_kept; ++r)
            for (int64_t k = src.A.ptr[kept[r]]; k < src.A.ptr[kept[r] + 1]; ++k)
                if (map[src.A.index[k]] >= 0) entries.push_back({r, map[src.A.index[k]], src.A.value[k]});
        m.A = Csr::from_entries(rows_kept, int64_t(m.c.size()), std::move(entries));
    }
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

double power_norm(const Csr &A, int iterations) {
    if (iterations < 1 || A.value.empty()) return 0;
    double scale = 0;
    for (double v : A.value) scale = std::max(scale, std::abs(v));
    if (!(scale > 0) || !std::isfinite(scale)) return 0;
    // On A / max|a|, so that A^T A cannot overflow from the units alone.
    Csr a = A;
    for (double &v : a.value) v /= scale;
    const Csr at = a.transposed();
    std::mt19937_64 random(0x56414e54414745ULL);
    std::vector<double> x(size_t(a.cols));
    for (double &v : x) v = double(random() >> 11) * 0x1p-53 - .5;
    // Euclidean norm by running hypot, in index order.
    auto norm_of = [](const std::vector<double> &v) {
        double n = 0;
        for (double e : v) n = std::hypot(n, e);
        return n;
    };
    auto normalize = [&](std::vector<double> &v) {
        const double n = norm_of(v);
        if (!(n > 0) || !std::isfinite(n)) return false;
        for (double &e : v) e /= n;
        return true;
    };
    if (!normalize(x)) return 0;
    double estimate = 0;
    std::vector<double> y(size_t(a.rows));
    for (int k = 0; k < iterations; ++k) {
        a.multiply(x.data(), y.data());
        estimate = std::max(estimate, norm_of(y));
        if (!normalize(y)) break;
        at.multiply(y.data(), x.data());
        if (!normalize(x)) break;
    }
    const double result = scale * estimate;
    return std::isfinite(result) ? result : 0;
}

double step_norm(const Model &m, const Options &o, double *estimate) {
    std::vector<double> colsum(size_t(m.cols()), 0);
    double rowmax = 0;
    for (int64_t i = 0; i < m.rows(); ++i) {
        double s = 0;
        for (int64_t k = m.A.ptr[i]; k < m.A.ptr[i + 1]; ++k) {
            double v = std::abs(m.A.value[k]);
            s += v;
            colsum[size_t(m.A.index[k])] += v;
        }
        rowmax = std::max(rowmax, s);
    }
    double colmax = 0;
    for (double s : colsum) colmax = std::max(colmax, s);
    double norm = std::sqrt(rowmax) * std::sqrt(colmax);
    if (estimate) *estimate = 0;
    if (o.power_iterations > 0) {
        const double e = power_norm(m.A, o.power_iterations);
        if (estimate) *estimate = e;
        // an estimate is not a bound: only backtracking steps may use it
        if (o.adaptive && e > 0) norm = std::min(norm, 1.05 * e);
    }
    return norm;
}

}  // namespace nyk
