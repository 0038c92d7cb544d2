// `niryukti` command-line front end of the MI355X PDHG engine.
//
// Commands, flags, output documents and exit codes follow Niryukti's CLI
// (app/main.cpp): solve (:478-652), verify (:429-477), inspect (:218-235),
// devices (:159-162) and session (app/session.hpp:4-18).  Exit codes: 0 for
// OPTIMAL / verified, 2 for any other solver outcome, 1 for errors.
#include <csignal>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>

#include "engine/io.hpp"
#include "engine/lp.hpp"
#include "engine/session.hpp"
#include "niryukti_pdhg.h"

namespace nyk {
extern volatile std::sig_atomic_t interrupted;
}

namespace {

void usage() {
    std::cout << "niryukti - restarted PDHG LP engine for AMD MI355X\n"
                 "usage: niryukti solve MODEL [options]\n"
                 "       niryukti verify MODEL RESULT.json\n"
                 "       niryukti inspect MODEL\n"
                 "       niryukti convert MODEL OUTPUT.mps|OUTPUT.json\n"
                 "       niryukti session MODEL        (line-JSON requests on stdin)\n"
                 "       niryukti devices\n"
                 "solve options: --device gpu|hip|auto --method pdhg|halpern|rhpdhg|r2hpdhg --tol X\n"
                 "       --time-limit S --iterations N --threads T --check-every K --scaling-passes P\n"
                 "       --scaling ruiz|combined --primal-weight displacement|pid --gpu-indices auto|32|64\n"
                 "       --matrix-precision fp64|mixed --device-prepare auto|on|off --power-iterations N\n"
                 "       --polishing --no-presolve --no-restart\n"
                 "       --no-adaptive --verbose\n"
                 "       --warm-start RESULT.json --allow-model-change --json-out FILE (--solution-out)\n"
                 "       --checkpoint-out FILE --resume FILE --checkpoint-nodes N\n"
                 "       --cuda-graphs and --gpu-monitor are accepted: graphs and device checks are always on\n"
                 "models: .mps or native .json\n";
}

int solve_command(int argc, char **argv, const nyk::Model &m) {
    nyk::Options o;
    o.method = "auto";
    std::string output, warm;
    bool allow_model_change = false, cuda_graphs = false, gpu_monitor = false;
    for (int i = 3; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&]() {
            if (++i >= argc) throw std::runtime_error("Missing option value for " + a);
            return std::string(argv[i]);
        };
        if (a == "--device") o.device = val();
        else if (a == "--tol") o.tol = std::stod(val());
        else if (a == "--time-limit") o.time_limit = std::stod(val());
        else if (a == "--iterations") o.iteration_limit = std::stoll(val());
        else if (a == "--threads") o.threads = std::stoi(val());
        else if (a == "--check-every") o.check_every = std::stoi(val());
        else if (a == "--scaling-passes") o.scaling_passes = std::stoi(val());
        else if (a == "--scaling") o.scaling = val();
        else if (a == "--method") o.method = val();
        else if (a == "--primal-weight") o.primal_weight = val();
        else if (a == "--gpu-indices") o.gpu_indices = val();
        else if (a == "--matrix-precision") o.matrix_precision = val();
        else if (a == "--power-iterations") o.power_iterations = std::stoi(val());
        else if (a == "--polishing") o.polishing = true;
        else if (a == "--device-prepare") {
            const std::string v = val();
            if (v != "auto" && v != "on" && v != "off") throw std::runtime_error("--device-prepare auto|on|off");
            o.device_prepare = v == "auto" ? -1 : v == "on" ? 1 : 0;
        }
        else if (a == "--json-out" || a == "--solution-out") output = val();
        else if (a == "--checkpoint-out") o.checkpoint_path = val();
        else if (a == "--resume") o.resume_path = val();
        else if (a == "--checkpoint-nodes") o.checkpoint_nodes = std::stoll(val());
        else if (a == "--warm-start") warm = val();
        else if (a == "--allow-model-change") allow_model_change = true;
        else if (a == "--verbose") o.verbose = true;
        else if (a == "--no-presolve") o.presolve = false;
        else if (a == "--no-restart") o.restart = false;
        else if (a == "--no-adaptive") o.adaptive = false;
        else if (a == "--cuda-graphs") cuda_graphs = true;
        else if (a == "--gpu-monitor") gpu_monitor = true;
        else if (a == "--auto") o.method = "auto";
        else
            throw std::runtime_error("Unknown option: " + a);
    }
    if (o.device == "auto") o.device = "gpu";
    if (!warm.empty()) {
        auto s = nyk::read_solution(m, warm, allow_model_change);
        o.initial_x = s.x;
        o.initial_y = s.y;
    }
    auto r = nyk::solve(m, o);
    auto j = nyk::result_json(m, r);
    // The options as requested (app/main.cpp:610-630; the branch-and-bound
    // keys of that block have no counterpart in an LP engine).
    auto &opt = j["options"];
    opt = nyk::Json::object();
    opt["tolerance"] = o.tol;
    opt["time_limit"] = o.time_limit;
    opt["iteration_limit"] = (long long)o.iteration_limit;
    opt["threads"] = o.threads;
    opt["scaling_passes"] = o.scaling_passes;
    opt["scaling"] = o.scaling;
    opt["method"] = o.method;
    opt["primal_weight"] = o.primal_weight;
    opt["power_iterations"] = o.power_iterations;
    opt["cuda_graphs"] = cuda_graphs;
    opt["gpu_monitor"] = gpu_monitor;
    opt["gpu_indices"] = o.gpu_indices;
    opt["matrix_precision"] = o.matrix_precision;
    opt["polishing"] = o.polishing;
    opt["restart"] = o.restart;
    opt["adaptive"] = o.adaptive;
    const std::string text = j.dump(2);
    if (!output.empty()) {
        std::ofstream f(output);
        if (!f) throw std::runtime_error("Cannot write " + output);
        f << text << '\n';
    }
    std::cout << text << '\n';
    return r.status == "OPTIMAL" ? 0 : 2;
}

// app/main.cpp:429-477
int verify_command(const nyk::Model &m, const std::string &path) {
    auto s = nyk::read_solution(m, path);
    nyk::Result r;
    r.backend = "verifier";
    r.method_selected = "verify";
    r.x = s.x;
    r.y = s.y;
    if (!s.infeasibility_ray.empty()) {
        nyk::Verifier verifier(m);
        r.certificate_margin = verifier.infeasibility_bound(s.infeasibility_ray);
        r.infeasibility_ray = s.infeasibility_ray;
        const bool pass = r.certificate_margin > 0 && std::isfinite(r.certificate_margin);
        r.status = pass ? "VERIFIED_INFEASIBLE" : "VERIFICATION_FAILED";
        std::cout << nyk::result_json(m, r).dump(2) << '\n';
        return pass ? 0 : 2;
    }
    r.accuracy = nyk::verify(m, r.x, r.y);
    bool pass = r.accuracy.finite && r.accuracy.primal <= 1e-6;
    bool optimal = pass && r.accuracy.kkt <= 1e-6;
    const double declared = s.objective * m.sense;  // canonical units
    const bool matches = !std::isfinite(s.objective) ||
                         (r.accuracy.finite &&
                          std::abs(declared - r.accuracy.objective) <= 1e-6 * (1 + std::abs(r.accuracy.objective)));
    if (!matches) {
        pass = optimal = false;
        r.message = "Reported objective does not match independent calculation";
    }
    if (s.status == "OPTIMAL" && !optimal) {
        pass = false;
        r.message = "Claimed optimal solution fails independent KKT checks";
    }
    r.status = optimal ? "VERIFIED_OPTIMAL" : pass ? "VERIFIED_FEASIBLE" : "VERIFICATION_FAILED";
    std::cout << nyk::result_json(m, r).dump(2) << '\n';
    return pass ? 0 : 2;
}

int session_command(const std::string &path) {
    nyk::Session session(path);
    nyk::Json ready = nyk::Json::object();
    ready["status"] = "READY";
    ready["fingerprint"] = session.fingerprint();
    std::cout << ready.dump() << std::endl;
    std::string line;
    while (std::getline(std::cin, line)) {
        try {
            auto req = nyk::Json::parse(line);
            const nyk::Json *action = req.find("action");
            if (action && action->is_string() && action->str() == "close") return 0;
            std::cout << session.request(req).dump() << std::endl;
        } catch (const std::exception &e) {
            nyk::Json err = nyk::Json::object();
            err["status"] = "ERROR";
            err["message"] = e.what();
            std::cout << err.dump() << std::endl;
        }
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    std::signal(SIGINT, [](int) { nyk::interrupted = 1; });
    try {
        if (argc < 2 || std::string(argv[1]) == "--help" || std::string(argv[1]) == "-h") {
            usage();
            return 0;
        }
        const std::string cmd = argv[1];
        if (cmd == "devices") {
            const int n = npdhg_device_count();
            std::cout << "CPU: not used for iterations (host presolve and certification only)\n";
            std::cout << "HIP: " << n << " device(s)\n";
            for (int d = 0; d < n; ++d) {
                char name[256] = {0};
                npdhg_device_name(d, name, sizeof name);
                std::cout << "  [" << d << "] " << name << '\n';
            }
            return 0;
        }
        if (argc < 3) throw std::runtime_error("Model path required");
        if (cmd == "session") return session_command(argv[2]);
        const nyk::Model m = nyk::read_model(argv[2]);
        if (cmd == "inspect") {
            nyk::Json j = nyk::Json::object();
            j["name"] = m.name;
            j["type"] = "LP";
            j["rows"] = (long long)m.rows();
            j["columns"] = (long long)m.cols();
            j["nonzeros"] = (long long)m.A.nnz();
            j["fingerprint"] = m.fingerprint();
            j["gpu_storage_estimate_bytes"] = 48. * m.A.nnz() + 240. * (m.rows() + m.cols() + 2);
            std::cout << j.dump(2) << '\n';
            return 0;
        }
        if (cmd == "verify") {
            if (argc != 4) throw std::runtime_error("verify MODEL SOLUTION");
            return verify_command(m, argv[3]);
        }
        if (cmd == "convert") {
            if (argc != 4) throw std::runtime_error("convert INPUT OUTPUT");
            nyk::write_model(m, argv[3]);
            return 0;
        }
        if (cmd == "solve") return solve_command(argc, argv, m);
        throw std::runtime_error("Unknown command: " + cmd);
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << '\n';
        return 1;
    }
}
