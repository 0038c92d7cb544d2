// Original-space KKT verification on the host: the certifying check.
//
// The check is laid out like the device evaluation (eval_kernel in
// hip/pdhg_backend.hip): one pass over the rows and one over the columns feed
// a small set of accumulators (maxima for the residuals, long-double sums for
// the objective support, complementarity and the box minimum of the
// Lagrangian), and a final step turns them into the rel-KKT figures.  The
// figures are the reference's (src/verify.cpp:103-240): normalised primal
// violations, the projected reduced-cost dual residual scaled by max(1+|c_j|),
// and the larger of the complementarity gap and the Lagrangian-bound gap.
// Sums run sequentially over rows, then columns, in long double, which is the
// order the reference's results depend on.  The certificate bound is an
// outward-rounded long-double interval evaluation of the Lagrangian dual
// function (src/verify.cpp:12-88), so it is a rigorous lower bound.
//
// The device evaluates the same quantities between certifications; this host
// version is the one that may declare OPTIMAL.
#include "lp.hpp"

#include <algorithm>

namespace nyk {
namespace {
using Wide = long double;
constexpr Wide kWideInf = std::numeric_limits<Wide>::infinity();

Wide lower_of(Wide v) { return std::nextafter(v, -kWideInf); }
Wide upper_of(Wide v) { return std::nextafter(v, kWideInf); }

// A long-double interval widened outward by one unit in the last place at
// every operation, so that it encloses the exact real value.
struct Enclosure {
    Wide lo, hi;
    void add_product(double a, double b) {
        const Wide p = Wide(a) * b;
        lo = lower_of(lo + lower_of(p));
        hi = upper_of(hi + upper_of(p));
    }
    bool holds_zero() const { return lo <= 0 && hi >= 0; }
};

// Smallest value of g * x over the box [l, u] for every g in `g`, rounded
// down; +inf when the box is unbounded in the direction g decreases.
Wide box_minimum(const Enclosure &g, double l, double u) {
    if (g.lo == 0 && g.hi == 0) return 0;
    Wide best = g.holds_zero() ? 0 : kWideInf;
    for (const double end : {l, u}) {
        if (!std::isfinite(end)) continue;
        for (const Wide slope : {g.lo, g.hi}) best = std::min(best, (end == 0 || slope == 0) ? 0 : lower_of(slope * end));
    }
    return best;
}

// Rigorous lower bound of the LP's Lagrangian dual function at y (with the
// objective and offset dropped in feasibility mode, for Farkas rays).  -inf
// when y pays for an infinite row bound or the columns' box cannot bound it.
double certified_lower_bound(const Model &m, const Csr &at, const std::vector<double> &y, bool feasibility) {
    Wide bound = feasibility ? 0 : m.offset;
    for (size_t i = 0; i < y.size(); ++i) {
        if (y[i] == 0) continue;
        const double side = y[i] > 0 ? m.ru[i] : m.rl[i];
        if (!std::isfinite(side)) return -kInf;
        if (side != 0) bound = lower_of(bound - upper_of(Wide(y[i]) * side));
    }
    for (int64_t j = 0; j < m.cols(); ++j) {
        const Wide cost = feasibility ? 0 : m.c[size_t(j)];
        Enclosure g{cost, cost};
        for (int64_t k = at.ptr[size_t(j)]; k < at.ptr[size_t(j) + 1]; ++k) {
            const double yi = y[size_t(at.index[size_t(k)])], a = at.value[size_t(k)];
            if (yi != 0 && a != 0) g.add_product(a, yi);
        }
        const double l = m.lb[size_t(j)], u = m.ub[size_t(j)];
        if ((!std::isfinite(l) && g.hi > 0) || (!std::isfinite(u) && g.lo < 0)) return -kInf;
        const Wide term = box_minimum(g, l, u);
        if (!std::isfinite(term)) return -kInf;
        if (term != 0) bound = lower_of(bound + term);
    }
    const double out = double(bound);
    return std::isfinite(out) ? std::nextafter(out, -kInf) : -kInf;
}

// Running figures of one evaluation.
struct Tally {
    double primal = 0, primal_abs = 0, dual = 0, dual_abs = 0;
    double objective = 0;
    Wide support = 0;     // sum of y_i times its active row bound
    Wide complement = 0;  // sum of |y_i (b_i - a_i x)| and |g_j (x_j - bound_j)|
    Wide box = 0;         // minimum of (c + A^T y)^T x over the box
    bool bounded = true;  // every active bound finite: the Lagrangian bound exists
    bool finite = true;
};

void tally_rows(const Model &m, const std::vector<double> &ax, const std::vector<double> &y, Tally &t) {
    for (size_t i = 0; i < y.size(); ++i) {
        const double lo = m.rl[i], hi = m.ru[i], row = ax[i];
        const double below = std::max(0., lo - row), above = std::max(0., row - hi);
        if (std::isfinite(lo)) t.primal = std::max(t.primal, below / (1 + std::abs(lo)));
        if (std::isfinite(hi)) t.primal = std::max(t.primal, above / (1 + std::abs(hi)));
        t.primal_abs = std::max(t.primal_abs, std::max(below, above));
        if (y[i] == 0) continue;
        const double side = y[i] > 0 ? hi : lo;
        if (std::isfinite(side)) {
            t.support += Wide(y[i]) * side;
            t.complement += std::abs(Wide(y[i]) * (side - row));
        } else {
            // a price on a side with no bound: dual infeasible by |y_i|
            t.bounded = false;
            t.dual = std::max(t.dual, std::abs(y[i]));
            t.dual_abs = std::max(t.dual_abs, std::abs(y[i]));
        }
    }
}

void tally_columns(const Model &m, const std::vector<double> &x, const std::vector<double> &aty, double cost_scale,
                   Tally &t) {
    for (size_t j = 0; j < x.size(); ++j) {
        const double xj = x[j], l = m.lb[j], u = m.ub[j];
        const double outside = std::max({0., l - xj, xj - u});
        t.primal = std::max(t.primal, outside / (1 + std::abs(xj)));
        t.primal_abs = std::max(t.primal_abs, outside);
        t.objective += m.c[j] * xj;
        const double reduced = m.c[j] + aty[j];  // gradient of the Lagrangian in x_j
        if (!std::isfinite(reduced)) {
            t.finite = false;
            return;
        }
        const double projected = std::clamp(reduced, xj - u, xj - l);
        t.dual = std::max(t.dual, std::abs(projected) / cost_scale);
        t.dual_abs = std::max(t.dual_abs, std::abs(projected));
        if (reduced > 0 && std::isfinite(l)) t.complement += std::abs(Wide(reduced) * (xj - l));
        if (reduced < 0 && std::isfinite(u)) t.complement += std::abs(Wide(reduced) * (u - xj));
        const double minimiser = reduced > 0 ? l : reduced < 0 ? u : std::clamp(0., l, u);
        if (std::isfinite(minimiser))
            t.box += Wide(reduced) * minimiser;
        else
            t.bounded = false;
    }
}

double relative_gap(double objective, double bound) {
    return std::abs(objective - bound) / (1 + std::abs(objective) + std::abs(bound));
}

bool all_finite(const std::vector<double> &v) {
    return std::all_of(v.begin(), v.end(), [](double e) { return std::isfinite(e); });
}
}  // namespace

Verifier::Verifier(const Model &model)
    : model_(model), transpose_(model.A.transposed()), activity_(model.A.rows), transpose_product_(model.A.cols) {}

double Verifier::infeasibility_bound(const std::vector<double> &y) const {
    if (y.size() != model_.rl.size() || !all_finite(y)) return -kInf;
    return certified_lower_bound(model_, transpose_, y, true);
}

Accuracy Verifier::evaluate(const std::vector<double> &x, const std::vector<double> &y, bool safe_bound) {
    const Model &m = model_;
    if (x.size() != m.c.size() || y.size() != m.rl.size() || !all_finite(x) || !all_finite(y)) return Accuracy{};
    m.A.multiply(x.data(), activity_.data());
    transpose_.multiply(y.data(), transpose_product_.data());
    if (!all_finite(activity_) || !all_finite(transpose_product_)) return Accuracy{};
    double cost_scale = 1;
    for (double c : m.c) cost_scale = std::max(cost_scale, 1 + std::abs(c));

    Tally t;
    t.objective = m.offset;
    tally_rows(m, activity_, y, t);
    tally_columns(m, x, transpose_product_, cost_scale, t);
    if (!t.finite) return Accuracy{};

    Accuracy a;
    a.objective = t.objective;
    a.primal = t.primal;
    a.primal_absolute = t.primal_abs;
    a.dual = t.dual;
    a.dual_absolute = t.dual_abs;
    a.complementarity = double(t.complement);
    a.gap = a.complementarity / (1 + std::abs(a.objective - m.offset));
    if (t.bounded) {
        // the Lagrangian value at y, lowered by a guard for its rounding
        const Wide value = Wide(m.offset) - t.support + t.box;
        const Wide guard = 64 * std::numeric_limits<double>::epsilon() *
                           (1 + std::abs(Wide(m.offset)) + std::abs(t.support) + std::abs(t.box));
        a.gap = std::max(a.gap, relative_gap(a.objective, double(value - guard)));
    }
    a.lower_bound = safe_bound ? certified_lower_bound(m, transpose_, y, false) : -kInf;
    if (std::isfinite(a.lower_bound)) a.gap = std::max(a.gap, relative_gap(a.objective, a.lower_bound));
    a.kkt = std::max({a.primal, a.dual, a.gap});
    a.finite = std::isfinite(a.objective) && std::isfinite(a.primal) && std::isfinite(a.dual) && std::isfinite(a.gap) &&
               std::isfinite(a.kkt);
    if (!a.finite) {
        a.kkt = kInf;
        a.lower_bound = -kInf;
    }
    return a;
}

Accuracy verify(const Model &m, const std::vector<double> &x, const std::vector<double> &y) {
    Verifier v(m);
    return v.evaluate(x, y);
}

}  // namespace nyk
