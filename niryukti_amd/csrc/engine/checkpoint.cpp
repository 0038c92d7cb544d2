// Continuous-solve checkpoint files (src/solver.cpp:384-509 defines the
// `vantage-continuous-1` schema: fingerprint, configuration, backend state,
// best pair, epoch, PID state, controller scalars and counters).  Non-finite
// values are stored as the strings "inf" / "-inf"; NaN is refused both ways.
#include "checkpoint.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace nyk {
namespace {
constexpr const char *kSchema = "vantage-continuous-1";

Json pack(const std::vector<double> &values) {
    Json out = Json::array();
    for (double v : values) {
        if (std::isnan(v)) throw std::runtime_error("NaN checkpoint state");
        if (std::isfinite(v))
            out.push_back(Json(v));
        else
            out.push_back(Json(v > 0 ? "inf" : "-inf"));
    }
    return out;
}

std::vector<double> unpack(const Json &array) {
    if (!array.is_array()) throw std::runtime_error("Invalid checkpoint array");
    std::vector<double> out;
    out.reserve(array.size());
    for (size_t k = 0; k < array.size(); ++k) {
        const Json &v = array[k];
        double q;
        if (v.is_number())
            q = v.number();
        else if (v.is_string() && v.str() == "inf")
            q = kInf;
        else if (v.is_string() && v.str() == "-inf")
            q = -kInf;
        else
            throw std::runtime_error("Invalid checkpoint number");
        if (std::isnan(q)) throw std::runtime_error("NaN checkpoint value");
        out.push_back(q);
    }
    return out;
}

int64_t count_of(const Json &doc, const char *key) {
    const Json &v = doc.at(key);
    if (!v.is_integral() || v.integer() < 0) throw std::runtime_error(std::string("Invalid checkpoint counter: ") + key);
    return v.integer();
}

Json read_document(const std::string &path) {
    std::ifstream file(path);
    if (!file) throw std::runtime_error("Cannot open continuous checkpoint");
    std::stringstream text;
    text << file.rdbuf();
    Json doc = Json::parse(text.str());
    if (!doc.is_object()) throw std::runtime_error("Invalid continuous checkpoint");
    return doc;
}

void check_model(const Json &doc, const Model &original) {
    const Json *schema = doc.find("schema");
    const Json *fp = doc.find("fingerprint");
    if (!schema || !schema->is_string() || schema->str() != kSchema || !fp || !fp->is_string() ||
        fp->str() != original.fingerprint())
        throw std::runtime_error("Continuous checkpoint model/configuration mismatch");
}
}  // namespace

Json checkpoint_configuration(const Options &o, bool device_checks) {
    Json c = Json::object();
    c["backend"] = "hip";
    c["method"] = o.method;
    c["tol"] = o.tol;
    c["presolve"] = o.presolve;
    c["scaling"] = o.scaling;
    c["scaling_passes"] = o.scaling_passes;
    c["adaptive"] = o.adaptive;
    c["restart"] = o.restart;
    c["primal_weight"] = o.primal_weight;
    c["power_iterations"] = o.power_iterations;
    c["polishing"] = o.polishing;
    c["check_every"] = o.check_every;
    c["graphs"] = o.graphs;
    c["precision"] = o.matrix_precision;
    c["indices"] = o.gpu_indices;
    c["device_checks"] = device_checks;
    return c;
}

void write_checkpoint(const std::string &path, const Model &original, const Json &configuration, const LoopState &s) {
    Json doc = Json::object();
    doc["schema"] = kSchema;
    doc["fingerprint"] = original.fingerprint();
    doc["configuration"] = configuration;
    Json state = Json::array();
    for (const auto &part : s.backend) state.push_back(pack(part));
    doc["backend_state"] = state;
    doc["best_x"] = pack(s.best_x);
    doc["best_y"] = pack(s.best_y);
    doc["epoch_x"] = pack(s.epoch_x);
    doc["epoch_y"] = pack(s.epoch_y);
    doc["pid"] = pack(s.pid);
    doc["control"] = pack({s.step, s.weight, s.restart_kkt, double(s.since_restart), double(s.next_polish),
                           double(s.previous_rejected), 0.0, 0.0});
    doc["iterations"] = (long long)s.iterations;
    doc["restarts"] = (long long)s.restarts;
    doc["rejected_steps"] = (long long)s.rejected_steps;
    doc["weight_updates"] = (long long)s.weight_updates;
    // this engine's additions: check counters, the decision trace and how the
    // best pair was rated
    Json extra = Json::object();
    extra["host_candidate_checks"] = (long long)s.host_checks;
    extra["device_candidate_checks"] = (long long)s.device_checks;
    extra["polishing_iterations"] = (long long)s.polishing_iterations;
    extra["polishing_attempts"] = (long long)s.polishing_attempts;
    extra["first_restart_iteration"] = (long long)s.first_restart_iteration;
    extra["first_restart_weight"] = pack({s.first_restart_weight});
    extra["best_certified"] = s.best_certified;
    extra["best_kkt"] = pack({s.best_kkt});
    extra["have_best"] = s.have_best;
    extra["fp32_phase"] = s.fp32_phase;
    extra["fp32_best"] = pack({s.fp32_best});
    extra["fp32_since"] = (long long)s.fp32_since;
    extra["fp32_iterations"] = (long long)s.fp32_iterations;
    doc["engine"] = extra;
    const std::string temporary = path + ".tmp";
    {
        std::ofstream file(temporary, std::ios::trunc);
        if (!file) throw std::runtime_error("Cannot write continuous checkpoint");
        file << doc.dump();
        file.flush();
        if (!file) throw std::runtime_error("Continuous checkpoint write failed");
    }
    if (std::rename(temporary.c_str(), path.c_str()) != 0) throw std::runtime_error("Continuous checkpoint rename failed");
}

void check_checkpoint_model(const std::string &path, const Model &original, const Json &configuration) {
    const Json doc = read_document(path);
    check_model(doc, original);
    // every setting but where the checks run, which is decided later
    const Json *config = doc.find("configuration");
    if (!config || !config->is_object()) throw std::runtime_error("Continuous checkpoint model/configuration mismatch");
    Json a = Json::object(), b = Json::object();
    for (const auto &kv : config->items())
        if (kv.first != "device_checks") a[kv.first] = kv.second;
    for (const auto &kv : configuration.items())
        if (kv.first != "device_checks") b[kv.first] = kv.second;
    if (a.dump() != b.dump()) throw std::runtime_error("Continuous checkpoint model/configuration mismatch");
}

LoopState read_checkpoint(const std::string &path, const Model &original, const Json &configuration) {
    const Json doc = read_document(path);
    check_model(doc, original);
    const Json *config = doc.find("configuration");
    if (!config || config->dump() != configuration.dump())
        throw std::runtime_error("Continuous checkpoint model/configuration mismatch");
    LoopState s;
    const Json &state = doc.at("backend_state");
    if (!state.is_array()) throw std::runtime_error("Invalid checkpoint backend state");
    for (size_t k = 0; k < state.size(); ++k) s.backend.push_back(unpack(state[k]));
    s.best_x = unpack(doc.at("best_x"));
    s.best_y = unpack(doc.at("best_y"));
    s.epoch_x = unpack(doc.at("epoch_x"));
    s.epoch_y = unpack(doc.at("epoch_y"));
    s.pid = unpack(doc.at("pid"));
    if (s.pid.size() != 3) throw std::runtime_error("PID checkpoint dimensions");
    const auto control = unpack(doc.at("control"));
    if (control.size() != 8) throw std::runtime_error("Continuous checkpoint controller dimensions");
    s.step = control[0];
    s.weight = control[1];
    s.restart_kkt = control[2];
    for (int k = 3; k <= 5; ++k)
        if (!(control[size_t(k)] >= 0) || control[size_t(k)] > 9.0e15 || control[size_t(k)] != std::floor(control[size_t(k)]))
            throw std::runtime_error("Invalid continuous checkpoint controller");
    s.since_restart = int64_t(control[3]);
    s.next_polish = int64_t(control[4]);
    s.previous_rejected = int64_t(control[5]);
    s.iterations = count_of(doc, "iterations");
    s.restarts = count_of(doc, "restarts");
    s.rejected_steps = count_of(doc, "rejected_steps");
    s.weight_updates = count_of(doc, "weight_updates");
    if (!(s.step > 0) || !(s.weight > 0) || !std::isfinite(s.step) || !std::isfinite(s.weight) ||
        std::isnan(s.restart_kkt))
        throw std::runtime_error("Invalid continuous checkpoint controller");
    const Json &extra = doc.at("engine");
    s.host_checks = count_of(extra, "host_candidate_checks");
    s.device_checks = count_of(extra, "device_candidate_checks");
    s.polishing_iterations = count_of(extra, "polishing_iterations");
    s.polishing_attempts = count_of(extra, "polishing_attempts");
    s.first_restart_iteration = count_of(extra, "first_restart_iteration");
    const auto frw = unpack(extra.at("first_restart_weight"));
    const auto bk = unpack(extra.at("best_kkt"));
    if (frw.size() != 1 || bk.size() != 1) throw std::runtime_error("Invalid continuous checkpoint extras");
    s.first_restart_weight = frw[0];
    s.best_kkt = bk[0];
    if (!extra.at("best_certified").is_bool() || !extra.at("have_best").is_bool())
        throw std::runtime_error("Invalid continuous checkpoint extras");
    s.best_certified = extra.at("best_certified").boolean();
    s.have_best = extra.at("have_best").boolean();
    if (!extra.at("fp32_phase").is_bool()) throw std::runtime_error("Invalid continuous checkpoint extras");
    s.fp32_phase = extra.at("fp32_phase").boolean();
    const auto fb = unpack(extra.at("fp32_best"));
    if (fb.size() != 1) throw std::runtime_error("Invalid continuous checkpoint extras");
    s.fp32_best = fb[0];
    s.fp32_since = count_of(extra, "fp32_since");
    s.fp32_iterations = count_of(extra, "fp32_iterations");
    if (s.best_x.size() != size_t(original.cols()) || s.best_y.size() != size_t(original.rows()))
        throw std::runtime_error("Continuous checkpoint solution dimensions");
    return s;
}

}  // namespace nyk
