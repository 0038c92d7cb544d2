// Original-space KKT verification on the host: the certifying check.
//
// Formulae follow the reference verifier (src/verify.cpp:12-240) term for term:
// per-row/per-column normalised primal residual, projected-gradient dual
// residual scaled by max(1 + |c_j|), complementarity and Lagrangian-bound gap,
// with long-double accumulation where the reference uses it and the
// outward-rounded long-double lower bound for the safe certificate.  The two
// SpMVs run in parallel by rows (thread-count independent); the accumulations
// stay sequential so the result is bit-identical to the reference.
//
// The GPU evaluates the same quantities between certifications; this host
// version is the one that may declare OPTIMAL.
#include "lp.hpp"

#include <algorithm>

namespace nyk {
namespace {
long double down(long double v) { return std::nextafter(v, -std::numeric_limits<long double>::infinity()); }
long double up(long double v) { return std::nextafter(v, std::numeric_limits<long double>::infinity()); }

// Outward-rounded Lagrangian lower bound (src/verify.cpp:12-88, LP form).
double lagrangian_lower_bound(const Model &m, const Csr &at, const std::vector<double> &y, bool feasibility) {
    long double lower = feasibility ? 0 : m.offset;
    for (size_t i = 0; i < y.size(); ++i)
        if (y[i] != 0) {
            double b = y[i] > 0 ? m.ru[i] : m.rl[i];
            if (!std::isfinite(b)) return -kInf;
            if (b == 0) continue;
            lower = down(lower - up((long double)y[i] * b));
        }
    for (size_t j = 0; j < m.c.size(); ++j) {
        long double gl = feasibility ? 0 : m.c[j], gu = gl;
        for (int64_t k = at.ptr[j]; k < at.ptr[j + 1]; ++k)
            if (y[at.index[k]] != 0 && at.value[k] != 0) {
                long double v = (long double)at.value[k] * y[at.index[k]];
                gl = down(gl + down(v));
                gu = up(gu + up(v));
            }
        if ((!std::isfinite(m.lb[j]) && gu > 0) || (!std::isfinite(m.ub[j]) && gl < 0)) return -kInf;
        long double term = std::numeric_limits<long double>::infinity();
        if (gl <= 0 && gu >= 0) term = 0;
        for (double b : {m.lb[j], m.ub[j]})
            if (std::isfinite(b)) {
                term = std::min(term, (b == 0 || gl == 0) ? 0 : down(gl * b));
                term = std::min(term, (b == 0 || gu == 0) ? 0 : down(gu * b));
            }
        if (gl == 0 && gu == 0) term = 0;
        if (!std::isfinite(term)) return -kInf;
        if (term != 0) lower = down(lower + term);
    }
    double result = double(lower);
    if (!std::isfinite(result)) return -kInf;
    return std::nextafter(result, -kInf);
}
}  // namespace

Verifier::Verifier(const Model &model)
    : model_(model), transpose_(model.A.transposed()), activity_(model.A.rows), transpose_product_(model.A.cols) {}

double Verifier::infeasibility_bound(const std::vector<double> &y) const {
    if (y.size() != model_.rl.size()) return -kInf;
    for (double v : y)
        if (!std::isfinite(v)) return -kInf;
    return lagrangian_lower_bound(model_, transpose_, y, true);
}

Accuracy Verifier::evaluate(const std::vector<double> &x, const std::vector<double> &y, bool safe_bound) {
    const Model &m = model_;
    Accuracy a;
    if (x.size() != m.c.size() || y.size() != m.rl.size()) return a;
    for (double v : x)
        if (!std::isfinite(v)) return a;
    for (double v : y)
        if (!std::isfinite(v)) return a;
    m.A.multiply(x.data(), activity_.data());
    transpose_.multiply(y.data(), transpose_product_.data());
    const auto &ax = activity_;
    const auto &aty = transpose_product_;
    for (double v : ax)
        if (!std::isfinite(v)) return Accuracy{};
    for (double v : aty)
        if (!std::isfinite(v)) return Accuracy{};
    a.objective = m.offset;
    a.primal = a.dual = a.primal_absolute = a.dual_absolute = 0;
    long double support = 0, box_min = 0, complement = 0;
    bool bound_ok = true;
    for (size_t i = 0; i < y.size(); ++i) {
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
    double cscale = 1;
    for (double v : m.c) cscale = std::max(cscale, 1 + std::abs(v));
    for (size_t j = 0; j < x.size(); ++j) {
        double v = std::max({0., m.lb[j] - x[j], x[j] - m.ub[j]});
        a.primal = std::max(a.primal, v / (1 + std::abs(x[j])));
        a.primal_absolute = std::max(a.primal_absolute, v);
        // The zero quadratic terms are kept so that overflow behaves as in the reference.
        a.objective += m.c[j] * x[j] + .5 * 0.0 * x[j] * x[j] + .5 * x[j] * 0.0;
        double linear = m.c[j] + aty[j], g = linear + 0.0 * x[j] + 0.0;
        if (!std::isfinite(linear) || !std::isfinite(g)) return Accuracy{};
        double d = std::clamp(g, x[j] - m.ub[j], x[j] - m.lb[j]);
        a.dual_absolute = std::max(a.dual_absolute, std::abs(d));
        a.dual = std::max(a.dual, std::abs(d) / cscale);
        if (g > 0 && std::isfinite(m.lb[j])) complement += std::abs((long double)g * (x[j] - m.lb[j]));
        if (g < 0 && std::isfinite(m.ub[j])) complement += std::abs((long double)g * (m.ub[j] - x[j]));
        double arg = linear > 0 ? m.lb[j] : linear < 0 ? m.ub[j] : std::clamp(0., m.lb[j], m.ub[j]);
        if (!std::isfinite(arg))
            bound_ok = false;
        else
            box_min += (long double)linear * arg + .5L * 0.0 * arg * arg;
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
    a.lower_bound = safe_bound ? lagrangian_lower_bound(m, transpose_, y, false) : -kInf;
    if (std::isfinite(a.lower_bound))
        a.gap = std::max(a.gap, std::abs(a.objective - a.lower_bound) / (1 + std::abs(a.objective) + std::abs(a.lower_bound)));
    a.kkt = std::max({a.primal, a.dual, a.gap});
    a.finite = std::isfinite(a.objective) && std::isfinite(a.kkt) && std::isfinite(a.gap) && std::isfinite(a.primal) &&
               std::isfinite(a.dual);
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
