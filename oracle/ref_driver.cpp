// Test-infrastructure driver around the reference library built by oracle/ref.mk.
// It generates the seeded instances in memory (no file parsing) or reads a model
// file with the reference's own reader, and calls the reference's public entry
// points: vantage::solve (src/solver.cpp:55), vantage::prepare
// (src/preprocess.cpp:133), vantage::cpu_backend (src/cpu.cpp:249) and
// vantage::verify (src/verify.cpp:241).  Output is one JSON object on stdout.
//
//   ref_driver solve   (--config C --seed S [--scale K] | --model FILE) [--threads T]
//                      [--tol X] [--iterations N] [--time-limit S] [--no-presolve]
//                      [--no-adaptive] [--no-restart] [--method M] [--xy]
//   ref_driver prepare (--config ... | --model FILE) [--no-presolve] [--passes P]
//   ref_driver steps   (--config ... | --model FILE) --count K --tau T --sigma S
//                      [--adaptive] [--halpern] [--reflection R] [--residual-restarts]
//                      [--reset-on-restart] [--chunks N]
//                      (--reset-on-restart: after a chunk that requested a restart,
//                      reset to the T(z) candidate, as the driver does)
//   ref_driver solve-range --config 2 --seed S --first I --stride K --budget SEC [--threads T]
//                      solves LPs I, I+K, I+2K, ... of the batched workload one after
//                      another until their summed solve() time reaches SEC; one JSON
//                      line per LP (instance generation is outside the timed solve)
//   ref_driver pid     --weight W --updates "dx,dy;dx,dy;..."   (PrimalWeightController,
//                      src/research.cpp:52-70: each update's weight feeds the next)
//
// This program is part of the checker only; nothing in the product links it.
#include "instances.hpp"
#include "internal.hpp"
#include "vantage/vantage.hpp"
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>

using namespace vantage;

namespace {
std::map<std::string, std::string> args;
bool flag(const char *k) { return args.count(k) > 0; }
std::string get(const char *k, const std::string &d) { return args.count(k) ? args[k] : d; }
double getd(const char *k, double d) { return args.count(k) ? std::stod(args[k]) : d; }
long long getl(const char *k, long long d) { return args.count(k) ? std::stoll(args[k]) : d; }

void num(double v) {
    if (std::isnan(v))
        std::printf("null");
    else if (std::isinf(v))
        std::printf(v > 0 ? "\"inf\"" : "\"-inf\"");
    else
        std::printf("%.17g", v);
}
void vec(const char *name, const std::vector<double> &v, bool comma = true) {
    std::printf("\"%s\":[", name);
    for (size_t i = 0; i < v.size(); ++i) {
        if (i) std::printf(",");
        num(v[i]);
    }
    std::printf("]%s", comma ? "," : "");
}
void ivec(const char *name, const std::vector<int64_t> &v, bool comma = true) {
    std::printf("\"%s\":[", name);
    for (size_t i = 0; i < v.size(); ++i) std::printf(i ? ",%lld" : "%lld", (long long)v[i]);
    std::printf("]%s", comma ? "," : "");
}

Model from_instance(const nyk_gen::Instance &p) {
    Model m;
    m.name = p.name;
    m.A.rows = p.rows;
    m.A.cols = p.cols;
    m.A.ptr = p.ptr;
    m.A.index = p.index;
    m.A.value = p.value;
    m.c = p.c;
    m.q.assign(p.cols, 0);
    m.lb = p.lb;
    m.ub = p.ub;
    m.rl = p.rl;
    m.ru = p.ru;
    m.types.assign(p.cols, VarType::Continuous);
    m.names.resize(p.cols);
    for (int64_t j = 0; j < p.cols; ++j) m.names[j] = "x" + std::to_string(j);
    m.row_names.resize(p.rows);
    for (int64_t i = 0; i < p.rows; ++i) m.row_names[i] = "r" + std::to_string(i);
    return m;
}

struct Loaded {
    Model model;
    bool planted = false;
    double planted_objective = 0;
};
Loaded load() {
    Loaded l;
    if (flag("--model")) {
        l.model = read_model(args["--model"]);
        return l;
    }
    int config = int(getl("--config", 0));
    uint64_t seed = uint64_t(getl("--seed", 42));
    nyk_gen::Instance p;
    if (config == 2)
        p = nyk_gen::batch_instance(nyk_gen::batch_shapes(int(getl("--index", 0)) + 1, seed).back());
    else if (config == 3)
        p = nyk_gen::batch_instance(nyk_gen::serve_shape(getl("--index", 0), seed));
    else if (config == 1 && flag("--alpha")) {
        int64_t sc = getl("--scale", 1);
        p = nyk_gen::powerlaw_lp(1000000 / sc, 2000000 / sc, seed, getd("--zipf", 1.5),
                                 int(getl("--mindeg", 2)), int(getl("--maxdeg", 25)), getl("--hubs", 64),
                                 getd("--alpha", 0.02), getd("--upper", 100), getd("--zero", 0.3),
                                 getd("--eq", 0.7));
    } else if (config == 5) {
        p = nyk_gen::equality_lp(getl("--m", 2000), getl("--n", 5000), int(getl("--lo", 5)),
                                 int(getl("--hi", 5)), getd("--upper", 10), seed, getd("--zero", 0.3));
    } else
        p = nyk_gen::workload(config, seed, getl("--scale", 1));
    l.model = from_instance(p);
    l.planted = true;
    l.planted_objective = p.objective;
    return l;
}

Options options() {
    Options o;
    o.device = "cpu";
    o.method = get("--method", "pdhg");
    o.threads = int(getl("--threads", 1));
    o.tol = getd("--tol", 1e-6);
    o.iteration_limit = getl("--iterations", 100000);
    o.time_limit = getd("--time-limit", 1e9);
    o.presolve = !flag("--no-presolve");
    o.adaptive = !flag("--no-adaptive");
    o.restart = !flag("--no-restart");
    o.scaling_passes = int(getl("--passes", 5));
    o.check_every = int(getl("--check-every", 100));
    o.scaling = get("--scaling", "ruiz");
    o.primal_weight = get("--primal-weight", "displacement");
    return o;
}

int cmd_solve() {
    auto l = load();
    auto o = options();
    auto t0 = std::chrono::steady_clock::now();
    auto r = solve(l.model, o);
    double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("{\"status\":\"%s\",", r.status.c_str());
    std::printf("\"objective\":");
    num(r.accuracy.finite ? l.model.sense * r.accuracy.objective : NAN);
    std::printf(",\"planted_objective\":");
    num(l.planted ? l.planted_objective : NAN);
    std::printf(",\"primal\":");
    num(r.accuracy.primal);
    std::printf(",\"dual\":");
    num(r.accuracy.dual);
    std::printf(",\"gap\":");
    num(r.accuracy.gap);
    std::printf(",\"kkt\":");
    num(r.accuracy.kkt);
    std::printf(",\"lower_bound\":");
    num(r.accuracy.lower_bound);
    std::printf(",\"complementarity\":");
    num(r.accuracy.complementarity);
    std::printf(",\"iterations\":%lld,\"restarts\":%lld,\"rejected_steps\":%lld,\"weight_updates\":%lld,"
                "\"host_candidate_checks\":%lld,\"removed_rows\":%lld,\"removed_columns\":%lld,",
                (long long)r.iterations, (long long)r.restarts, (long long)r.rejected_steps,
                (long long)r.weight_updates, (long long)r.host_candidate_checks, (long long)r.removed_rows,
                (long long)r.removed_columns);
    std::printf("\"seconds\":%.9g,\"iteration_seconds\":%.9g,\"preprocess_seconds\":%.9g,"
                "\"verification_seconds\":%.9g,\"wall_seconds\":%.9g,\"threads\":%d,",
                r.seconds, r.iteration_seconds, r.preprocess_seconds, r.verification_seconds, wall,
                o.threads);
    std::printf("\"rows\":%lld,\"cols\":%lld,\"nnz\":%lld,\"fingerprint\":\"%s\"",
                (long long)l.model.A.rows, (long long)l.model.A.cols,
                (long long)l.model.A.value.size(), l.model.fingerprint().c_str());
    if (flag("--xy")) {
        std::printf(",");
        vec("x", r.x);
        vec("y", r.y, false);
    }
    std::printf("}\n");
    return 0;
}

int cmd_prepare() {
    auto l = load();
    auto o = options();
    auto p = prepare(l.model, o);
    const auto &m = p.model;
    std::printf("{\"failure\":\"%s\",\"objective_scale\":", p.failure.c_str());
    num(p.objective_scale);
    std::printf(",\"offset\":");
    num(m.offset);
    std::printf(",\"rows\":%lld,\"cols\":%lld,", (long long)m.A.rows, (long long)m.A.cols);
    ivec("col_map", p.cols);
    ivec("row_map", p.rows);
    ivec("lower_rows", p.lower_rows);
    ivec("upper_rows", p.upper_rows);
    vec("fixed", p.fixed);
    vec("column_scale", p.column_scale);
    vec("row_scale", p.row_scale);
    vec("c", m.c);
    vec("lb", m.lb);
    vec("ub", m.ub);
    vec("rl", m.rl);
    vec("ru", m.ru);
    ivec("ptr", m.A.ptr);
    ivec("index", m.A.index);
    vec("value", m.A.value, false);
    std::printf("}\n");
    return 0;
}

// Raw backend trajectory on the generated (unscaled) model: chunks of `count`
// advance() calls with fixed tau/sigma from x0 = clamp(0, lb, ub), y0 = 0,
// candidates recorded after each chunk.
int cmd_steps() {
    auto l = load();
    const auto &m = l.model;
    std::vector<double> x(m.c.size()), y(m.rl.size());
    for (size_t j = 0; j < x.size(); ++j) x[j] = std::clamp(0.0, m.lb[j], m.ub[j]);
    auto b = cpu_backend(m, x, y, flag("--adaptive"), flag("--halpern"), getd("--reflection", 0),
                         flag("--residual-restarts"));
    int count = int(getl("--count", 10)), chunks = int(getl("--chunks", 3));
    double tau = getd("--tau", 0.1), sigma = getd("--sigma", 0.1);
    std::printf("{\"chunks\":[");
    for (int k = 0; k < chunks; ++k) {
        int acc = b->advance(count, tau, sigma);
        std::vector<double> cx, cy, ax, ay;
        b->candidates(cx, cy, ax, ay);
        std::printf("%s{\"accepted\":%d,\"rejected\":%lld,\"restart_requested\":%d,", k ? "," : "", acc,
                    (long long)b->rejected_steps(), b->restart_requested() ? 1 : 0);
        vec("x", cx);
        vec("y", cy);
        vec("xa", ax);
        vec("ya", ay, false);
        std::printf("}");
        if (flag("--reset-on-restart") && b->restart_requested()) b->reset(ax, ay);
    }
    std::printf("]}\n");
    return 0;
}
int cmd_solve_range() {
    const uint64_t seed = uint64_t(getl("--seed", 42));
    const int64_t first = getl("--first", 0), stride = std::max<int64_t>(1, getl("--stride", 1));
    const int64_t total = getl("--total", 4096);
    const double budget = getd("--budget", 10);
    auto o = options();
    const auto shapes = nyk_gen::batch_shapes(int(total), seed);
    double spent = 0;
    for (int64_t i = first; i < total && spent < budget; i += stride) {
        Model m = from_instance(nyk_gen::batch_instance(shapes[size_t(i)]));
        auto t0 = std::chrono::steady_clock::now();
        auto r = solve(m, o);
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        spent += dt;
        std::printf("{\"index\":%lld,\"status\":\"%s\",\"iterations\":%lld,\"seconds\":%.9g,\"kkt\":",
                    (long long)i, r.status.c_str(), (long long)r.iterations, dt);
        num(r.accuracy.kkt);
        std::printf("}\n");
        std::fflush(stdout);
    }
    return 0;
}

// The reference's PID primal-weight controller on a sequence of epoch
// displacements; prints the weight after each update.
int cmd_pid() {
    PrimalWeightController pid;
    double w = getd("--weight", 1);
    std::string seq = get("--updates", "");
    std::printf("{\"weights\":[");
    size_t at = 0;
    bool first = true;
    while (at < seq.size()) {
        size_t end = seq.find(';', at);
        if (end == std::string::npos) end = seq.size();
        std::string item = seq.substr(at, end - at);
        size_t comma = item.find(',');
        if (comma == std::string::npos) throw std::runtime_error("pid: updates are dx,dy pairs");
        long double dx = std::stold(item.substr(0, comma)), dy = std::stold(item.substr(comma + 1));
        w = pid.update(w, (long double)(double)dx, (long double)(double)dy);
        std::printf(first ? "" : ",");
        num(w);
        first = false;
        at = end + 1;
    }
    std::printf("]}\n");
    return 0;
}
}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: ref_driver solve|prepare|steps [options]\n");
        return 1;
    }
    for (int i = 2; i < argc; ++i) {
        std::string k = argv[i];
        if (k.rfind("--", 0) != 0) continue;
        bool boolean = i + 1 >= argc || std::strncmp(argv[i + 1], "--", 2) == 0;
        if (k == "--scale" || k == "--seed" || k == "--config" || k == "--tau" || k == "--sigma" ||
            k == "--reflection" || k == "--alpha" || k == "--zipf" || k == "--maxdeg" ||
            k == "--upper" || k == "--zero" || k == "--m" || k == "--n" || k == "--lo" ||
            k == "--hi" || k == "--eq" || k == "--mindeg" || k == "--scaling" || k == "--hubs" ||
            k == "--weight" || k == "--updates" || k == "--primal-weight" || k == "--method")
            boolean = i + 1 >= argc;
        args[k] = boolean ? "1" : argv[++i];
    }
    try {
        std::string c = argv[1];
        if (c == "solve") return cmd_solve();
        if (c == "prepare") return cmd_prepare();
        if (c == "steps") return cmd_steps();
        if (c == "pid") return cmd_pid();
        if (c == "solve-range") return cmd_solve_range();
        std::fprintf(stderr, "unknown command\n");
        return 1;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
