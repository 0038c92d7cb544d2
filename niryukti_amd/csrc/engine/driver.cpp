// Restarted-PDHG driver: presolve/scale on the host, iterate on the GPU,
// certify in the original space on the host.
//
// Control flow and every decision rule follow the reference's continuous
// driver (src/solver.cpp:215-735, PDHG branch): chunked advance of
// `check_every` accepted steps, candidate check of the current and averaged
// iterates in original units, best-candidate tracking, OPTIMAL only after a
// safe re-evaluation, Farkas ray tests after 1000 iterations, the restart rule
// and the displacement / PID primal-weight update.  Iterations run in the HIP
// backend behind the npdhg_* C-ABI (include/niryukti_pdhg.h).
#include "checkpoint.hpp"
#include "lp.hpp"
#include "niryukti_pdhg.h"
#include "niryukti_prepare.h"
#include "parallel.hpp"

#include <algorithm>
#include <chrono>
#include <csignal>
#include <cstdio>
#include <stdexcept>

namespace nyk {

volatile std::sig_atomic_t interrupted = 0;

namespace {
using Clock = std::chrono::steady_clock;
double elapsed(Clock::time_point s) { return std::chrono::duration<double>(Clock::now() - s).count(); }

bool anchored(const Options &o) { return o.method == "halpern" || o.method == "rhpdhg" || o.method == "r2hpdhg"; }
bool stop_requested(const Options &o) {
    return interrupted || (o.cancellation && o.cancellation->load(std::memory_order_relaxed));
}

// PID controller on the log primal weight (src/research.cpp:52-70).
class WeightPid {
    double integral_ = 0, previous_ = 0;
    bool initialized_ = false;

  public:
    double update(double weight, long double dx, long double dy) {
        if (!(weight > 0) || !std::isfinite(weight) || !(dx > 1e-24L) || !(dy > 1e-24L) || !std::isfinite(dx) ||
            !std::isfinite(dy))
            return weight;
        double error = double(std::log((long double)weight) + .5L * (std::log(dx) - std::log(dy)));
        error = std::clamp(error, -10., 10.);
        double next_integral = std::clamp(integral_ + error, -100., 100.);
        double derivative = initialized_ ? error - previous_ : 0;
        double correction = std::clamp(.3 * error + .01 * next_integral + .1 * derivative, -std::log(4.), std::log(4.));
        double logarithm = std::log(weight) - correction;
        double clipped = std::clamp(logarithm, std::log(1e-4), std::log(1e4));
        if (clipped == logarithm) integral_ = next_integral;
        previous_ = error;
        initialized_ = true;
        return std::exp(clipped);
    }
    std::vector<double> state() const { return {integral_, previous_, initialized_ ? 1. : 0.}; }
    void restore(const std::vector<double> &v) {
        if (v.size() != 3 || !std::isfinite(v[0]) || !std::isfinite(v[1])) throw std::runtime_error("PID checkpoint state");
        integral_ = v[0];
        previous_ = v[1];
        initialized_ = v[2] != 0;
    }
};

// --matrix-precision fp32: the attempts start in FP32 arithmetic (FP32
// values, gathered vectors and per-lane sums; FP64 iterates, steps, checks
// and decisions) and move to FP64 for the rest of the solve, the refinement
// phase, at the first check whose candidate reaches kSwitchKkt (about a
// hundred times the rel-KKT resolution of FP32 sums) or after the FP32 phase
// has gone kStallChecks checks without halving its best rel-KKT.
class Refinement {
  public:
    static constexpr double kSwitchKkt = 1e-5;
    static constexpr int kStallChecks = 20;
    explicit Refinement(const Options &o) : active_(o.matrix_precision == "fp32") {}
    // After a check with candidate rel-KKT `kkt`: true when the refinement phase starts now.
    bool switch_now(double kkt) {
        if (!active_) return false;
        if (kkt < 0.5 * best_) {
            best_ = kkt;
            since_ = 0;
        } else {
            ++since_;
        }
        if (kkt <= kSwitchKkt || since_ >= kStallChecks) {
            active_ = false;
            return true;
        }
        return false;
    }
    bool in_fp32() const { return active_; }
    void restore(bool fp32, double best, int since) {
        active_ = fp32;
        best_ = best;
        since_ = since;
    }
    double best() const { return best_; }
    int since() const { return since_; }

  private:
    bool active_;
    double best_ = kInf;
    int since_ = 0;
};

// RAII owner of a device backend.
class Device {
  public:
    // `at`, when not empty, is the transpose of m.A (from the device prepare).
    Device(const Model &m, const std::vector<double> &x, const std::vector<double> &y, const npdhg_config &cfg,
           const Csr &at) {
        npdhg_lp lp{m.rows(), m.cols(), m.A.ptr.data(), m.A.index.data(), m.A.value.data(), m.c.data(),
                    m.lb.data(),  m.ub.data(), m.rl.data(),      m.ru.data()};
        if (at.ptr.empty())
            check(npdhg_create(&lp, x.data(), y.data(), &cfg, &b_));
        else
            check(npdhg_create_transposed(&lp, at.ptr.data(), at.index.data(), at.value.data(), x.data(), y.data(),
                                          &cfg, &b_));
    }
    ~Device() { npdhg_destroy(b_); }
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;
    int advance(int count, double tau, double sigma) {
        int accepted = 0;
        check(npdhg_advance(b_, count, tau, sigma, &accepted));
        return accepted;
    }
    npdhg_status status() {
        npdhg_status s{};
        check(npdhg_status_get(b_, &s));
        return s;
    }
    void candidates(std::vector<double> &x, std::vector<double> &y, std::vector<double> &xa, std::vector<double> &ya) {
        check(npdhg_candidates(b_, x.data(), y.data(), xa.data(), ya.data()));
    }
    void reset_candidate(bool average) { check(npdhg_reset_candidate(b_, average ? 1 : 0)); }
    void reset(const std::vector<double> &x, const std::vector<double> &y) { check(npdhg_reset(b_, x.data(), y.data())); }

    // Device-side checks against a resident copy of the original model; false
    // when the prepare maps cannot be inverted (the driver then checks on the host).
    bool attach(const Model &original, const Prepared &prep) {
        npdhg_original o{};
        o.rows = original.rows();
        o.cols = original.cols();
        o.ptr = original.A.ptr.data();
        o.index = original.A.index.data();
        o.value = original.A.value.data();
        o.c = original.c.data();
        o.lb = original.lb.data();
        o.ub = original.ub.data();
        o.rl = original.rl.data();
        o.ru = original.ru.data();
        o.offset = original.offset;
        o.cols_map = prep.cols.data();
        o.lower_rows = prep.lower_rows.data();
        o.upper_rows = prep.upper_rows.data();
        o.fixed = prep.fixed.data();
        o.column_scale = prep.column_scale.data();
        o.row_scale = prep.row_scale.data();
        o.objective_scale = prep.objective_scale;
        return npdhg_attach_original(b_, &o) == 0;
    }
    void evaluate(npdhg_accuracy &current, npdhg_accuracy &average) { check(npdhg_evaluate(b_, &current, &average)); }
    void keep_best(bool average) { check(npdhg_keep_best(b_, average ? 1 : 0)); }
    void original_point(int which, std::vector<double> *x, std::vector<double> *y) {
        check(npdhg_original_point(b_, which, x ? x->data() : nullptr, y ? y->data() : nullptr));
    }
    bool farkas_displacement(bool average, std::vector<double> &ray) {
        int possible = 0;
        check(npdhg_farkas_displacement(b_, average ? 1 : 0, &possible, ray.data()));
        return possible != 0;
    }
    void epoch_displacement(bool average, long double &dx, long double &dy) {
        double a = 0, b = 0;
        check(npdhg_epoch_displacement(b_, average ? 1 : 0, &a, &b));
        dx = a;
        dy = b;
    }
    void set_epoch(bool average) { check(npdhg_set_epoch(b_, average ? 1 : 0)); }
    void leave_fp32_phase() { check(npdhg_set_fp32_phase(b_, 0)); }
    std::vector<double> snapshot() {
        int64_t size = 0;
        check(npdhg_snapshot_size(b_, &size));
        std::vector<double> s(static_cast<size_t>(size));
        check(npdhg_snapshot(b_, s.data()));
        return s;
    }
    void restore(const std::vector<double> &s) {
        int64_t size = 0;
        check(npdhg_snapshot_size(b_, &size));
        if (s.size() != size_t(size)) throw std::runtime_error("Continuous checkpoint backend dimensions");
        check(npdhg_restore(b_, s.data()));
    }

  private:
    static void check(int rc) {
        if (rc != 0) throw std::runtime_error(npdhg_last_error());
    }
    npdhg_backend *b_ = nullptr;
};

std::string device_description() {
    char name[256] = "HIP device";
    if (npdhg_device_count() > 0) npdhg_device_name(0, name, sizeof name);
    return name;
}

constexpr int64_t kDeviceCheckNonzeros = 100000;
constexpr int64_t kDevicePrepareNonzeros = 100000;

Accuracy from_device(const npdhg_accuracy &d) {
    Accuracy a;
    a.finite = d.finite != 0;
    if (!a.finite) return a;
    a.objective = d.objective;
    a.primal = d.primal;
    a.dual = d.dual;
    a.gap = d.gap;
    a.kkt = d.kkt;
    return a;
}

// Checkpoint files of one solve: the backend snapshot is stored as its
// arrays (iterates, averages, activity, anchors, scalars, and with device
// checks the epoch and best points), as the reference stores its backend's.
class Checkpointer {
  public:
    Checkpointer(const Options &o, const Model &original, Device &backend, int64_t n, int64_t m, bool device_checks)
        : o_(o), original_(original), backend_(backend), config_(checkpoint_configuration(o, device_checks)) {
        parts_ = {n, m, n, m, m, n, m, m, 8};
        if (device_checks) parts_.insert(parts_.end(), {n, m, original.cols(), original.rows(), 1});
        if (o.checkpoint_nodes < 1) throw std::runtime_error("checkpoint_nodes must be at least 1");
    }
    bool writing() const { return !o_.checkpoint_path.empty(); }
    bool resuming() const { return !o_.resume_path.empty(); }
    // At the top of a chunk: write when the iteration count has reached the next mark.
    bool due(int64_t iterations) {
        if (!writing() || iterations < next_) return false;
        next_ = iterations + o_.checkpoint_nodes;
        return true;
    }
    void save(LoopState &s) {
        const auto flat = backend_.snapshot();
        s.backend.clear();
        size_t at = 0;
        for (int64_t len : parts_) {
            s.backend.emplace_back(flat.begin() + long(at), flat.begin() + long(at + size_t(len)));
            at += size_t(len);
        }
        write_checkpoint(o_.checkpoint_path, original_, config_, s);
    }
    // Reads the resume file and restores the backend from it.
    LoopState load() {
        LoopState s = read_checkpoint(o_.resume_path, original_, config_);
        if (s.backend.size() != parts_.size()) throw std::runtime_error("Continuous checkpoint backend state count");
        std::vector<double> flat;
        for (size_t k = 0; k < parts_.size(); ++k) {
            if (s.backend[k].size() != size_t(parts_[k]))
                throw std::runtime_error("Continuous checkpoint backend dimensions");
            flat.insert(flat.end(), s.backend[k].begin(), s.backend[k].end());
        }
        backend_.restore(flat);
        next_ = s.iterations;
        return s;
    }

  private:
    const Options &o_;
    const Model &original_;
    Device &backend_;
    Json config_;
    std::vector<int64_t> parts_;
    int64_t next_ = 0;
};

// Host-side Farkas test of one ray (src/solver.cpp:599-633): normalise by the
// largest entry, then the exact outward-rounded bound decides.
bool certify_ray(Verifier &verifier, std::vector<double> ray, Result &r) {
    double norm_y = 0;
    for (double v : ray) {
        if (!std::isfinite(v)) return false;
        norm_y = std::max(norm_y, std::abs(v));
    }
    if (!(norm_y > 0)) return false;
    for (double &v : ray) v /= norm_y;
    const double margin = verifier.infeasibility_bound(ray);
    if (!(margin > 0 && std::isfinite(margin))) return false;
    r.status = "INFEASIBLE";
    r.message = "Independently verified box/row Farkas certificate";
    r.infeasibility_ray = std::move(ray);
    r.certificate_margin = margin;
    return true;
}

Result solve_continuous(const Model &original, const Options &o);

// The dual feasibility problem of an LP (src/research.cpp:72-94): variables y
// (one per row, sign-restricted where a row bound is infinite) and one
// stationarity row per column, c_j + (A^T y)_j free or one-signed by the
// column's bounds, zero objective.
Model dual_feasibility_model(const Model &m) {
    Model d;
    d.name = m.name + "_dual_feasibility";
    d.A = m.A.transposed();
    d.c.assign(size_t(m.rows()), 0);
    d.offset = 0;
    d.sense = 1;
    for (int64_t i = 0; i < m.rows(); ++i) {
        d.names.push_back("y" + std::to_string(i));
        d.lb.push_back(std::isfinite(m.rl[i]) ? -kInf : 0);
        d.ub.push_back(std::isfinite(m.ru[i]) ? kInf : 0);
    }
    for (int64_t j = 0; j < m.cols(); ++j) {
        d.row_names.push_back("stationarity" + std::to_string(j));
        d.rl.push_back(std::isfinite(m.ub[j]) ? -kInf : -m.c[j]);
        d.ru.push_back(std::isfinite(m.lb[j]) ? kInf : -m.c[j]);
    }
    d.validate();
    return d;
}

// Feasibility polishing (src/solver.cpp:634-690).  From iteration 100 on, at
// the first check past each doubling of the iteration count, when the chosen
// candidate's gap is <= 1 %: solve the zero-objective primal feasibility
// problem from its x, then the dual feasibility problem from its y, each with
// up to 1/8 of the iterations so far; the pair (x, y) they return is kept when
// the verifier rates it better.  The sub-solves' work counts in r.
struct Polish {
    int64_t next = 100;
    bool due(const Options &o, const Result &r, const Accuracy &ca, Clock::time_point start) {
        if (!o.polishing || r.iterations < next) return false;
        next = r.iterations <= std::numeric_limits<int64_t>::max() / 2 ? 2 * r.iterations
                                                                         : std::numeric_limits<int64_t>::max();
        return ca.gap <= .01 && budget(o, r) > 0 && !stop_requested(o) && elapsed(start) < o.time_limit;
    }
    static int64_t budget(const Options &o, const Result &r) {
        return std::min(r.iterations / 8, (o.iteration_limit - r.iterations) / 2);
    }
    // Returns true (with x, y, accuracy) when both sub-solves produced a pair.
    static bool run(const Model &original, const Options &o, Verifier &verifier, const std::vector<double> &ox,
                    const std::vector<double> &oy, Result &r, Clock::time_point start, std::vector<double> &x,
                    std::vector<double> &y, Accuracy &accuracy) {
        const int64_t limit = budget(o, r);
        r.polishing_attempts++;
        auto accumulate = [&](const Result &aux) {
            r.iterations += aux.iterations;
            r.polishing_iterations += aux.iterations;
            r.restarts += aux.restarts;
            r.rejected_steps += aux.rejected_steps;
            r.weight_updates += aux.weight_updates;
            r.preprocess_seconds += aux.preprocess_seconds;
            r.transfer_seconds += aux.transfer_seconds;
            r.iteration_seconds += aux.iteration_seconds;
            r.verification_seconds += aux.verification_seconds;
        };
        Model primal = original;
        std::fill(primal.c.begin(), primal.c.end(), 0.0);
        primal.offset = 0;
        Options po = o;
        po.checkpoint_path.clear();  // the sub-solves are not checkpointed
        po.resume_path.clear();
        po.method = "pdhg";
        po.adaptive = true;
        po.polishing = false;
        po.initial_x = ox;
        po.initial_y.assign(size_t(original.rows()), 0);
        po.iteration_limit = limit;
        po.time_limit = std::max(0., o.time_limit - elapsed(start));
        Result pr = solve_continuous(primal, po);
        accumulate(pr);
        if (!(pr.accuracy.finite && pr.accuracy.primal <= o.tol) || stop_requested(o) ||
            elapsed(start) >= o.time_limit)
            return false;
        const Model dual = dual_feasibility_model(original);
        po.initial_x = oy;
        po.initial_y.assign(size_t(original.cols()), 0);
        po.iteration_limit = std::min(limit, o.iteration_limit - r.iterations);
        po.time_limit = std::max(0., o.time_limit - elapsed(start));
        Result dr = solve_continuous(dual, po);
        accumulate(dr);
        if (dr.x.size() != size_t(original.rows())) return false;
        accuracy = verifier.evaluate(pr.x, dr.x);
        x = std::move(pr.x);
        y = std::move(dr.x);
        return true;
    }
};

// The chunk loop of solve_continuous with the candidate checks on the device:
// the same decisions in the same order (src/solver.cpp:540-726), fed by
// npdhg_evaluate's original-space KKT figures.  The host downloads a point
// only to certify OPTIMAL with its own verifier and safe bound, to test a
// Farkas ray the device could not rule out, and once at the end for the best
// candidate.
void run_device_checks(const Model &original, const Prepared &prep, const Options &o, Verifier &verifier,
                       Device &backend, Result &r, Clock::time_point start, double step, double weight,
                       double restart_kkt, WeightPid &controller, Checkpointer &ck) {
    const bool residual_restarts = o.method == "rhpdhg" || o.method == "r2hpdhg";
    const int64_t n0 = original.cols(), m0 = original.rows();
    int64_t since_restart = 0, previous_rejected = 0;
    double best_kkt = r.accuracy.kkt;
    bool have_best = false;
    Polish polish;
    Refinement refine(o);
    if (ck.resuming()) {
        LoopState s = ck.load();
        r.x = std::move(s.best_x);
        r.y = std::move(s.best_y);
        r.accuracy = verifier.evaluate(r.x, r.y);
        step = s.step;
        weight = s.weight;
        restart_kkt = s.restart_kkt;
        since_restart = s.since_restart;
        previous_rejected = s.previous_rejected;
        polish.next = s.next_polish;
        best_kkt = s.best_kkt;
        have_best = s.have_best;
        refine.restore(s.fp32_phase, s.fp32_best, int(s.fp32_since));
        if (o.matrix_precision == "fp32" && !s.fp32_phase) backend.leave_fp32_phase();
        r.fp32_iterations = s.fp32_iterations;
        controller.restore(s.pid);
        r.iterations = s.iterations;
        r.restarts = s.restarts;
        r.rejected_steps = s.rejected_steps;
        r.weight_updates = s.weight_updates;
        r.host_candidate_checks = s.host_checks;
        r.device_candidate_checks = s.device_checks;
        r.polishing_iterations = s.polishing_iterations;
        r.polishing_attempts = s.polishing_attempts;
        r.first_restart_iteration = s.first_restart_iteration;
        r.first_restart_weight = s.first_restart_weight;
    }
    while (r.iterations < o.iteration_limit) {
        if (ck.due(r.iterations)) {
            LoopState s;
            s.best_x = r.x;
            s.best_y = r.y;
            s.pid = controller.state();
            s.step = step;
            s.weight = weight;
            s.restart_kkt = restart_kkt;
            s.since_restart = since_restart;
            s.next_polish = polish.next;
            s.previous_rejected = previous_rejected;
            s.iterations = r.iterations;
            s.restarts = r.restarts;
            s.rejected_steps = r.rejected_steps;
            s.weight_updates = r.weight_updates;
            s.host_checks = r.host_candidate_checks;
            s.device_checks = r.device_candidate_checks;
            s.polishing_iterations = r.polishing_iterations;
            s.polishing_attempts = r.polishing_attempts;
            s.first_restart_iteration = r.first_restart_iteration;
            s.first_restart_weight = r.first_restart_weight;
            s.best_kkt = best_kkt;
            s.have_best = have_best;
            s.fp32_phase = refine.in_fp32();
            s.fp32_best = refine.best();
            s.fp32_since = refine.since();
            s.fp32_iterations = r.fp32_iterations;
            ck.save(s);
        }
        if (stop_requested(o)) {
            r.status = "INTERRUPTED";
            break;
        }
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
        auto st = backend.status();
        r.rejected_steps += st.rejected_total - previous_rejected;
        previous_rejected = st.rejected_total;
        const bool restart_requested = st.restart_requested != 0;
        r.device_candidate_checks++;
        tick = Clock::now();
        npdhg_accuracy dc{}, da{};
        backend.evaluate(dc, da);
        Accuracy ca = from_device(dc);
        const Accuracy aa = from_device(da);
        const bool avg = aa.kkt < ca.kkt;
        const bool chosen_ray = avg ? da.farkas_possible : dc.farkas_possible;
        const bool other_ray = avg ? dc.farkas_possible : da.farkas_possible;
        if (avg) ca = aa;
        if (ca.finite && ca.kkt < best_kkt) {
            backend.keep_best(avg);
            best_kkt = ca.kkt;
            have_best = true;
        }
        r.verification_seconds += elapsed(tick);
        if (refine.in_fp32()) r.fp32_iterations = r.iterations;
        if (refine.switch_now(ca.kkt)) backend.leave_fp32_phase();
        if (o.verbose)
            std::fprintf(stderr, "iter=%lld objective=%.12g primal=%.3e dual=%.3e gap=%.3e seconds=%.3f\n",
                         (long long)r.iterations, original.sense * ca.objective, ca.primal, ca.dual, ca.gap,
                         elapsed(start));
        if (!ca.finite) {
            r.status = "NUMERICAL_ERROR";
            r.message = "Nonfinite iterate or diagnostic";
            break;
        }
        if (ca.kkt <= o.tol) {
            tick = Clock::now();
            r.host_candidate_checks++;
            std::vector<double> ox(n0), oy(m0);
            backend.original_point(avg ? 1 : 0, &ox, &oy);
            auto safe = verifier.evaluate(ox, oy);
            r.verification_seconds += elapsed(tick);
            if (safe.kkt <= o.tol) {
                r.x = std::move(ox);
                r.y = std::move(oy);
                r.accuracy = safe;
                r.status = "OPTIMAL";
                return;
            }
            ca = safe;
        }
        if (r.iterations >= 1000 && ca.primal > o.tol) {
            std::vector<double> ray(m0);
            bool found = false;
            if (chosen_ray) {
                backend.original_point(avg ? 1 : 0, nullptr, &ray);
                found = certify_ray(verifier, ray, r);
            }
            if (!found && other_ray) {
                backend.original_point(avg ? 0 : 1, nullptr, &ray);
                found = certify_ray(verifier, ray, r);
            }
            if (!found && backend.farkas_displacement(avg, ray)) found = certify_ray(verifier, ray, r);
            if (found) break;
        }
        if (polish.due(o, r, ca, start)) {
            std::vector<double> ox(n0), oy(m0), px, py;
            backend.original_point(avg ? 1 : 0, &ox, &oy);
            Accuracy pa;
            if (Polish::run(original, o, verifier, ox, oy, r, start, px, py, pa) && pa.finite && pa.kkt < best_kkt) {
                r.x = std::move(px);
                r.y = std::move(py);
                r.accuracy = pa;
                best_kkt = pa.kkt;
                have_best = false;  // r holds the best pair until the device finds a better one
                if (pa.kkt <= o.tol) {
                    r.status = "OPTIMAL";
                    return;
                }
            }
        }
        const bool restart_due = residual_restarts ? restart_requested
                                                   : since_restart >= int64_t(o.check_every) * 2 &&
                                                         (ca.kkt < .5 * restart_kkt ||
                                                          since_restart >= std::max<int64_t>(2000, r.iterations / 2));
        if (o.restart && restart_due) {
            const bool to_average = residual_restarts || avg;
            if (o.adaptive || anchored(o) || o.primal_weight == "pid") {  // solver.cpp:702-717
                long double dx = 0, dy = 0;
                backend.epoch_displacement(to_average, dx, dy);
                if (dx > 1e-24 && dy > 1e-24) {
                    if (o.primal_weight == "pid")
                        weight = controller.update(weight, dx, dy);
                    else
                        weight = std::sqrt(weight * std::clamp(std::sqrt(double(dy / dx)), 1e-4, 1e4));
                    r.weight_updates++;
                }
            }
            backend.reset_candidate(to_average);
            backend.set_epoch(false);
            restart_kkt = ca.kkt;
            since_restart = 0;
            if (r.restarts++ == 0) {
                r.first_restart_iteration = r.iterations;
                r.first_restart_weight = weight;
            }
        }
    }
    auto check = Clock::now();
    if (have_best) backend.original_point(2, &r.x, &r.y);
    r.accuracy = verifier.evaluate(r.x, r.y);
    r.verification_seconds += elapsed(check);
}

// Ruiz passes, objective scale and transpose on the device for a prepared
// (presolved, unscaled) LP: fills the scaled model and factors into `p`
// exactly as the host passes would (src/preprocess.cpp:306-345) and returns
// the transpose of the scaled matrix.
Csr scale_on_device(Prepared &p, int passes) {
    Model &m = p.model;
    npdhg_lp lp{m.rows(), m.cols(), m.A.ptr.data(), m.A.index.data(), m.A.value.data(), m.c.data(),
                m.lb.data(),  m.ub.data(), m.rl.data(),      m.ru.data()};
    Csr at;
    at.rows = m.cols();
    at.cols = m.rows();
    at.ptr.resize(size_t(m.cols() + 1));
    at.index.resize(m.A.value.size());
    at.value.resize(m.A.value.size());
    std::vector<double> value(m.A.value.size()), c(m.c.size()), lb(m.c.size()), ub(m.c.size()), rl(m.rl.size()),
        ru(m.rl.size());
    npdhg_scaled out{};
    out.value = value.data();
    out.c = c.data();
    out.lb = lb.data();
    out.ub = ub.data();
    out.rl = rl.data();
    out.ru = ru.data();
    out.row_scale = p.row_scale.data();
    out.column_scale = p.column_scale.data();
    out.t_ptr = at.ptr.data();
    out.t_index = at.index.data();
    out.t_value = at.value.data();
    if (npdhg_scale_lp(&lp, passes, &out) != 0)
        throw std::runtime_error(std::string("device prepare: ") + npdhg_scale_last_error());
    m.A.value = std::move(value);
    m.c = std::move(c);
    m.lb = std::move(lb);
    m.ub = std::move(ub);
    m.rl = std::move(rl);
    m.ru = std::move(ru);
    if (passes > 0) {
        p.objective_scale = out.objective_scale;
        m.offset /= p.objective_scale;
    }
    m.validate();
    return at;
}

bool wants_device_prepare(const Model &original, const Options &o) {
    if (o.scaling != "ruiz" || original.A.nnz() >= (int64_t(1) << 31)) return false;
    return o.device_prepare > 0 || (o.device_prepare < 0 && original.A.nnz() >= kDevicePrepareNonzeros);
}

Result solve_continuous(const Model &original, const Options &o) {
    auto start = Clock::now();
    Result r;
    r.method_selected = o.method;
    r.backend = "hip";
    r.estimated_gpu_bytes = 48. * original.A.nnz() + 240. * (original.rows() + original.cols() + 2);
    r.device_reason = "Explicit backend request";
    Verifier verifier(original);
    r.status = "ITERATION_LIMIT";
    const bool device_prepare = wants_device_prepare(original, o);
    Options host_part = o;
    if (device_prepare) host_part.scaling_passes = 0;  // presolve only; the passes run on the device
    auto prep = prepare(original, host_part);
    Csr transposed;
    if (device_prepare && prep.failure.empty()) {
        transposed = scale_on_device(prep, o.scaling_passes);
        r.device_prepare = true;
    }
    r.preprocess_seconds = elapsed(start);
    if (!prep.failure.empty()) {
        r.status = prep.failure;
        r.message = prep.reason;
        r.infeasibility_ray = prep.failure_ray;
        if (!r.infeasibility_ray.empty()) r.certificate_margin = verifier.infeasibility_bound(r.infeasibility_ray);
        r.seconds = elapsed(start);
        return r;
    }
    const Model &m = prep.model;
    r.removed_columns = original.cols() - m.cols();
    r.removed_rows = original.rows() - m.rows();
    std::vector<double> x(m.cols()), y(m.rows());
    for (size_t j = 0; j < x.size(); ++j) {
        double initial = o.initial_x.empty() ? 0 : o.initial_x[prep.cols[j]];
        x[j] = std::clamp(initial / prep.column_scale[j], m.lb[j], m.ub[j]);
    }
    if (!o.initial_y.empty())
        for (size_t i = 0; i < y.size(); ++i) y[i] = o.initial_y[prep.rows[i]] / (prep.row_scale[i] * prep.objective_scale);
    r.x = prep.restore_x(x);
    r.y = prep.restore_y(y, original.rl.size());
    r.accuracy = verifier.evaluate(r.x, r.y);
    // Feasible start plus an isolated improving column proves unboundedness (solver.cpp:269-288).
    if (r.accuracy.finite && r.accuracy.primal_absolute == 0) {
        std::vector<uint8_t> used(m.cols(), 0);
        for (auto j : m.A.index) used[j] = 1;
        for (int64_t j = 0; j < m.cols(); ++j)
            if (!used[j] && ((m.c[j] < 0 && m.ub[j] == kInf) || (m.c[j] > 0 && m.lb[j] == -kInf))) {
                r.status = "UNBOUNDED";
                r.message = "Feasible point and isolated improving recession direction: " + m.column_name(j);
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
    // ||A||_2 <= sqrt(||A||_1 ||A||_inf) gives the guaranteed step (solver.cpp:299-322).
    const double norm = step_norm(m, o, &r.operator_norm_estimate);
    double step = norm > 0 ? .9 / norm : 1, weight = 1;
    if ((o.adaptive || anchored(o)) && o.primal_weight != "pid") {  // solver.cpp:323-333
        long double c2 = 0, b2 = 0;
        for (double c : m.c) c2 += (long double)c * c;
        for (int64_t i = 0; i < m.rows(); ++i) {
            double b = std::clamp(0., m.rl[i], m.ru[i]);
            b2 += (long double)b * b;
        }
        if (c2 > 1e-24 && b2 > 1e-24) weight = std::clamp(std::sqrt(double(c2 / b2)), 1e-4, 1e4);
    }
    r.initial_step = step;
    r.initial_weight = weight;
    const bool residual_restarts = o.method == "rhpdhg" || o.method == "r2hpdhg";
    auto transfer = Clock::now();
    npdhg_config cfg{};
    cfg.adaptive = o.adaptive;
    cfg.halpern = anchored(o);
    cfg.reflection = o.method == "r2hpdhg" ? 1 : 0;
    cfg.residual_restarts = residual_restarts && o.restart;
    cfg.graphs = o.graphs;
    cfg.index_bits = o.gpu_indices == "64" ? 64 : o.gpu_indices == "32" ? 32 : 0;
    cfg.fp32_values = o.matrix_precision == "mixed";
    cfg.fp32_arithmetic = o.matrix_precision == "fp32";
    Device backend(m, x, y, cfg, transposed);
    transposed = Csr{};
    r.transfer_seconds = elapsed(transfer);
    r.device_name = device_description();
    r.gpu_index_bits = backend.status().index_bits;
    r.graph_execution = o.graphs;
    r.matrix_precision = o.matrix_precision;

    double restart_kkt = r.accuracy.kkt;
    int64_t since_restart = 0, previous_rejected = 0;
    WeightPid controller;
    const bool want_device = o.device_checks > 0 || (o.device_checks < 0 && original.A.nnz() >= kDeviceCheckNonzeros);
    const bool device_checks = want_device && backend.attach(original, prep);
    Checkpointer ck(o, original, backend, m.cols(), m.rows(), device_checks);
    if (device_checks) {
        run_device_checks(original, prep, o, verifier, backend, r, start, step, weight, restart_kkt, controller, ck);
        r.seconds = elapsed(start);
        return r;
    }
    std::vector<double> epochx = x, epochy = y;
    std::vector<double> cx(m.cols()), cy(m.rows()), ax(m.cols()), ay(m.rows());
    Polish polish;
    Refinement refine(o);
    bool best_certified = true;  // r holds the start point's figures, certified bound included
    if (ck.resuming()) {
        LoopState s = ck.load();
        if (s.epoch_x.size() != size_t(m.cols()) || s.epoch_y.size() != size_t(m.rows()))
            throw std::runtime_error("Continuous checkpoint epoch dimensions");
        r.x = std::move(s.best_x);
        r.y = std::move(s.best_y);
        best_certified = s.best_certified;
        r.accuracy = verifier.evaluate(r.x, r.y, best_certified);
        if (!r.accuracy.finite) throw std::runtime_error("Invalid continuous checkpoint candidate");
        epochx = std::move(s.epoch_x);
        epochy = std::move(s.epoch_y);
        step = s.step;
        weight = s.weight;
        restart_kkt = s.restart_kkt;
        since_restart = s.since_restart;
        previous_rejected = s.previous_rejected;
        polish.next = s.next_polish;
        refine.restore(s.fp32_phase, s.fp32_best, int(s.fp32_since));
        if (o.matrix_precision == "fp32" && !s.fp32_phase) backend.leave_fp32_phase();
        r.fp32_iterations = s.fp32_iterations;
        controller.restore(s.pid);
        r.iterations = s.iterations;
        r.restarts = s.restarts;
        r.rejected_steps = s.rejected_steps;
        r.weight_updates = s.weight_updates;
        r.host_candidate_checks = s.host_checks;
        r.device_candidate_checks = s.device_checks;
        r.polishing_iterations = s.polishing_iterations;
        r.polishing_attempts = s.polishing_attempts;
        r.first_restart_iteration = s.first_restart_iteration;
        r.first_restart_weight = s.first_restart_weight;
    }
    while (r.iterations < o.iteration_limit) {
        if (ck.due(r.iterations)) {
            LoopState s;
            s.best_x = r.x;
            s.best_y = r.y;
            s.best_certified = best_certified;
            s.epoch_x = epochx;
            s.epoch_y = epochy;
            s.pid = controller.state();
            s.step = step;
            s.weight = weight;
            s.restart_kkt = restart_kkt;
            s.since_restart = since_restart;
            s.next_polish = polish.next;
            s.previous_rejected = previous_rejected;
            s.iterations = r.iterations;
            s.restarts = r.restarts;
            s.rejected_steps = r.rejected_steps;
            s.weight_updates = r.weight_updates;
            s.host_checks = r.host_candidate_checks;
            s.device_checks = r.device_candidate_checks;
            s.polishing_iterations = r.polishing_iterations;
            s.polishing_attempts = r.polishing_attempts;
            s.first_restart_iteration = r.first_restart_iteration;
            s.first_restart_weight = r.first_restart_weight;
            s.fp32_phase = refine.in_fp32();
            s.fp32_best = refine.best();
            s.fp32_since = refine.since();
            s.fp32_iterations = r.fp32_iterations;
            ck.save(s);
        }
        if (stop_requested(o)) {
            r.status = "INTERRUPTED";
            break;
        }
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
        auto st = backend.status();
        r.rejected_steps += st.rejected_total - previous_rejected;
        previous_rejected = st.rejected_total;
        const bool restart_requested = st.restart_requested != 0;
        r.host_candidate_checks++;
        tick = Clock::now();
        backend.candidates(cx, cy, ax, ay);
        std::vector<double> restart_x, restart_y;
        if (residual_restarts && restart_requested) {
            restart_x = ax;
            restart_y = ay;
        }
        auto ox = prep.restore_x(cx);
        auto oy = prep.restore_y(cy, original.rl.size());
        auto ca = verifier.evaluate(ox, oy, false);
        auto aox = prep.restore_x(ax);
        auto aoy = prep.restore_y(ay, original.rl.size());
        auto aa = verifier.evaluate(aox, aoy, false);
        const bool avg = aa.kkt < ca.kkt;
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
            best_certified = false;
        }
        r.verification_seconds += elapsed(tick);
        if (refine.in_fp32()) r.fp32_iterations = r.iterations;
        if (refine.switch_now(ca.kkt)) backend.leave_fp32_phase();
        if (o.verbose)
            std::fprintf(stderr, "iter=%lld objective=%.12g primal=%.3e dual=%.3e gap=%.3e seconds=%.3f\n",
                         (long long)r.iterations, original.sense * ca.objective, ca.primal, ca.dual, ca.gap,
                         elapsed(start));
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
        // Farkas rays from the iterates and the epoch displacement (solver.cpp:599-633).
        if (r.iterations >= 1000 && ca.primal > o.tol) {
            auto displacement = cy;
            for (size_t i = 0; i < displacement.size(); ++i) displacement[i] -= epochy[i];
            std::vector<std::vector<double>> rays{oy, aoy, prep.restore_y(displacement, original.rl.size())};
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
        if (polish.due(o, r, ca, start)) {
            std::vector<double> px, py;
            Accuracy pa;
            if (Polish::run(original, o, verifier, ox, oy, r, start, px, py, pa) && pa.finite &&
                pa.kkt < r.accuracy.kkt) {
                r.x = std::move(px);
                r.y = std::move(py);
                r.accuracy = pa;
                best_certified = true;
                if (pa.kkt <= o.tol) {
                    r.status = "OPTIMAL";
                    break;
                }
            }
        }
        const bool restart_due = residual_restarts ? restart_requested
                                                   : since_restart >= int64_t(o.check_every) * 2 &&
                                                         (ca.kkt < .5 * restart_kkt ||
                                                          since_restart >= std::max<int64_t>(2000, r.iterations / 2));
        if (o.restart && restart_due) {
            if (residual_restarts) {
                cx = std::move(restart_x);
                cy = std::move(restart_y);
            }
            if (o.adaptive || anchored(o) || o.primal_weight == "pid") {  // solver.cpp:702-717
                long double dx = 0, dy = 0;
                for (size_t j = 0; j < cx.size(); ++j) dx += (long double)(cx[j] - epochx[j]) * (cx[j] - epochx[j]);
                for (size_t i = 0; i < cy.size(); ++i) dy += (long double)(cy[i] - epochy[i]) * (cy[i] - epochy[i]);
                if (dx > 1e-24 && dy > 1e-24) {
                    if (o.primal_weight == "pid")
                        weight = controller.update(weight, dx, dy);
                    else
                        weight = std::sqrt(weight * std::clamp(std::sqrt(double(dy / dx)), 1e-4, 1e4));
                    r.weight_updates++;
                }
            }
            backend.reset_candidate(residual_restarts || avg);
            epochx = cx;
            epochy = cy;
            restart_kkt = ca.kkt;
            since_restart = 0;
            if (r.restarts++ == 0) {
                r.first_restart_iteration = r.iterations;
                r.first_restart_weight = weight;
            }
        }
        cx.resize(m.cols());
        cy.resize(m.rows());
        ax.resize(m.cols());
        ay.resize(m.rows());
    }
    auto check = Clock::now();
    r.accuracy = verifier.evaluate(r.x, r.y);
    r.verification_seconds += elapsed(check);
    if (r.status == "OPTIMAL" && (!r.accuracy.finite || r.accuracy.kkt > o.tol)) r.status = "NUMERICAL_ERROR";
    r.seconds = elapsed(start);
    return r;
}
}  // namespace

// src/solver.cpp:18-54, restricted to the options this engine implements.
void validate_options(const Options &o) {
    if (!(o.tol > 0 && o.tol < 1) || !std::isfinite(o.tol) || o.iteration_limit < 0 || o.time_limit < 0 ||
        std::isnan(o.time_limit) || o.check_every < 1 || o.scaling_passes < 0 || o.scaling_passes > 20 || o.threads < 1 ||
        o.power_iterations < 0 || o.power_iterations > 1000)
        throw std::runtime_error("Invalid solver options");
    if (o.device != "auto" && o.device != "cpu" && o.device != "cuda" && o.device != "hip" && o.device != "gpu")
        throw std::runtime_error("Unknown device");
    if (o.method != "pdhg" && o.method != "auto" && !anchored(o))
        throw std::runtime_error("Unknown method: use pdhg, halpern, rhpdhg or r2hpdhg");
    if (o.primal_weight != "displacement" && o.primal_weight != "pid")
        throw std::runtime_error("Unknown primal weight controller");
    if (o.scaling != "ruiz" && o.scaling != "combined") throw std::runtime_error("Unknown scaling: use ruiz or combined");
    if (o.gpu_indices != "auto" && o.gpu_indices != "32" && o.gpu_indices != "64")
        throw std::runtime_error("GPU indices must be auto, 32 or 64");
    if (o.matrix_precision != "fp64" && o.matrix_precision != "mixed" && o.matrix_precision != "fp32")
        throw std::runtime_error("Matrix precision must be fp64, mixed or fp32");
    if (anchored(o) && o.matrix_precision != "fp64") throw std::runtime_error("Fixed Halpern operators require FP64 matrices");
}

// src/solver.cpp:55-105 for continuous LPs.
Result solve(const Model &m, const Options &options) {
    auto overall = Clock::now();
    m.validate();
    validate_options(options);
    Options o = options;
    if (o.method == "auto") o.method = "pdhg";
    auto requested = [&](Result &r) {
        r.method_requested = options.method;
        r.automatic = options.method == "auto";
        r.scaling = o.scaling;
        r.tol = o.tol;
        r.time_limit = o.time_limit;
        r.iteration_limit = o.iteration_limit;
        r.scaling_passes = o.scaling_passes;
        if (r.method_selected.empty()) r.method_selected = o.method;
    };
    if (o.device == "cpu" || o.device == "cuda") {
        Result r;
        requested(r);
        r.status = "UNSUPPORTED";
        r.backend = o.device;
        r.message = "Requested backend not compiled: " + o.device + " (this engine runs PDHG on HIP; use --device gpu)";
        return r;
    }
    if (anchored(o) && o.adaptive) throw std::runtime_error("Experimental Halpern supports continuous LP with --no-adaptive");
    if (!o.initial_x.empty() && o.initial_x.size() != size_t(m.cols()))
        throw std::runtime_error("Warm-start primal dimension mismatch");
    if (!o.initial_y.empty() && o.initial_y.size() != size_t(m.rows()))
        throw std::runtime_error("Warm-start dual dimension mismatch");
    for (double v : o.initial_x)
        if (!std::isfinite(v)) throw std::runtime_error("Nonfinite warm start");
    for (double v : o.initial_y)
        if (!std::isfinite(v)) throw std::runtime_error("Nonfinite warm start");
    if (!o.resume_path.empty()) check_checkpoint_model(o.resume_path, m, checkpoint_configuration(o, false));
    if (npdhg_device_count() < 1) {
        Result r;
        requested(r);
        r.status = "UNSUPPORTED";
        r.message = "No HIP device is visible";
        return r;
    }
    auto result = solve_continuous(m, o);
    requested(result);
    result.seconds = elapsed(overall);
    return result;
}

}  // namespace nyk
