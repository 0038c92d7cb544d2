// Command-line front end of the CPU oracle (test infrastructure and the
// benchmark's CPU-baseline leg only).  Generates a seeded workload in memory
// and solves it with the oracle's FP64 restated PDHG; prints one JSON object.
//
//   pdhg_oracle --config C --seed S [--scale K] [--index I] [--threads T] [--tol X]
//               [--iterations N] [--time-limit S] [--scaling ruiz|combined] [--method M]
#include "instances.hpp"
#include "pdhg_oracle.h"
#include <cstdio>
#include <cstring>
#include <map>
#include <string>

int main(int argc, char **argv) {
    std::map<std::string, std::string> a;
    for (int i = 1; i + 1 < argc; i += 2) a[argv[i]] = argv[i + 1];
    auto get = [&](const char *k, const char *d) { return a.count(k) ? a[k] : std::string(d); };
    int config = std::stoi(get("--config", "0"));
    uint64_t seed = std::stoull(get("--seed", "42"));
    nyk_gen::Instance p;
    try {
        if (config == 2)
            p = nyk_gen::batch_instance(nyk_gen::batch_shapes(std::stoi(get("--index", "0")) + 1, seed).back());
        else if (config == 3)
            p = nyk_gen::batch_instance(nyk_gen::serve_shape(std::stoll(get("--index", "0")), seed));
        else
            p = nyk_gen::workload(config, seed, std::stoll(get("--scale", "1")));
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    oracle_lp lp{p.rows, p.cols, p.ptr.data(), p.index.data(), p.value.data(), p.c.data(), p.lb.data(),
                 p.ub.data(), p.rl.data(), p.ru.data(), 0.0};
    oracle_options o;
    oracle_default_options(&o);
    std::string scaling = get("--scaling", "ruiz"), method = get("--method", "pdhg");
    o.scaling = scaling.c_str();
    o.method = method.c_str();
    o.threads = std::stoi(get("--threads", "1"));
    o.tol = std::stod(get("--tol", "1e-6"));
    o.iteration_limit = std::stoll(get("--iterations", "100000"));
    o.time_limit = std::stod(get("--time-limit", "1e9"));
    oracle_result r;
    std::memset(&r, 0, sizeof r);
    char err[512] = {0};
    if (oracle_solve(&lp, &o, &r, err, sizeof err) != 0) {
        std::fprintf(stderr, "error: %s\n", err);
        return 1;
    }
    std::printf("{\"status\":\"%s\",\"objective\":%.17g,\"planted_objective\":%.17g,\"kkt\":%.17g,"
                "\"primal\":%.17g,\"dual\":%.17g,\"gap\":%.17g,\"iterations\":%lld,\"restarts\":%lld,"
                "\"rejected_steps\":%lld,\"weight_updates\":%lld,\"seconds\":%.9g,\"iteration_seconds\":%.9g,"
                "\"threads\":%d,\"rows\":%lld,\"cols\":%lld,\"nnz\":%lld}\n",
                r.status, r.accuracy.objective, p.objective, r.accuracy.kkt, r.accuracy.primal, r.accuracy.dual,
                r.accuracy.gap, (long long)r.iterations, (long long)r.restarts, (long long)r.rejected_steps,
                (long long)r.weight_updates, r.seconds, r.iteration_seconds, o.threads, (long long)p.rows,
                (long long)p.cols, (long long)p.nnz());
    return 0;
}
