// Model session (update / solve with warm start) and the vantage_* / niryukti_*
// C-ABI.  Protocol and error behaviour follow Niryukti's Session
// (src/session.hpp:12-73) and C API (src/c_api.cpp:11-46); the one deliberate
// difference is locking: each handle has its own mutex instead of one
// process-wide lock, so independent sessions can solve concurrently.
#include "session.hpp"

#include <algorithm>
#include <cmath>
#include <csignal>
#include <cstdio>
#include <mutex>
#include <stdexcept>
#include <unordered_map>

#include "niryukti/c_api.h"
#include "vantage/c_api.h"

namespace nyk {

extern volatile std::sig_atomic_t interrupted;

Session::Session(const std::string &path) : model_(read_model(path)) {}

Json Session::request(const Json &req) {
    const std::string action = req.at("action").str();
    if (action == "verify_certificate")
        throw std::runtime_error("UNSUPPORTED: this engine does not issue verification certificates");
    const std::vector<std::string> allowed =
        action == "update" ? std::vector<std::string>{"action", "rows", "variables", "objective"}
                           : std::vector<std::string>{"action", "device", "method", "time_limit", "tol",
                                                      "threads", "iterations", "warm_start"};
    for (const auto &kv : req.items())
        if (std::find(allowed.begin(), allowed.end(), kv.first) == allowed.end())
            throw std::runtime_error("Unknown session setting: " + kv.first);
    if (action == "update") return update(req);
    if (action == "solve") return solve(req);
    throw std::runtime_error("Unknown session action");
}

// src/session.hpp:22-51
Json Session::update(const Json &req) {
    Model candidate = model_;
    auto index_of = [&](bool rows) {
        std::unordered_map<std::string, int64_t> ids;
        const int64_t n = rows ? model_.rows() : model_.cols();
        ids.reserve(size_t(n));
        for (int64_t k = 0; k < n; ++k) ids.emplace(rows ? model_.row_name(k) : model_.column_name(k), k);
        return ids;
    };
    for (const char *section : {"rows", "variables", "objective"}) {
        const Json *sec = req.find(section);
        if (!sec) continue;
        if (!sec->is_object()) throw std::runtime_error("Updates must be named objects");
        const bool rows = std::string(section) == "rows";
        const auto ids = index_of(rows);
        for (const auto &kv : sec->items()) {
            auto it = ids.find(kv.first);
            if (it == ids.end()) throw std::runtime_error("Unknown model name: " + kv.first);
            const int64_t k = it->second;
            if (std::string(section) == "objective") {
                const double value = kv.second.number();
                if (!std::isfinite(value)) throw std::runtime_error("Nonfinite objective");
                candidate.c[k] = model_.sense * value;
                continue;
            }
            if (!kv.second.is_object()) throw std::runtime_error("Bounds must be an object");
            for (const auto &b : kv.second.items()) {
                if (b.first != "lb" && b.first != "ub") throw std::runtime_error("Expected lb or ub");
                const bool lower = b.first == "lb";
                const double value = b.second.is_null() ? (lower ? -kInf : kInf) : b.second.number();
                if (!b.second.is_null() && !std::isfinite(value)) throw std::runtime_error("Nonfinite bound");
                auto &values = rows ? (lower ? candidate.rl : candidate.ru) : (lower ? candidate.lb : candidate.ub);
                values[k] = value;
            }
        }
    }
    candidate.validate();
    model_ = std::move(candidate);
    Json out = Json::object();
    out["status"] = "UPDATED";
    out["fingerprint"] = model_.fingerprint();
    return out;
}

// src/session.hpp:52-72
Json Session::solve(const Json &req) {
    Options o;
    auto str = [&](const char *key, const char *fallback) {
        const Json *j = req.find(key);
        return j ? j->str() : std::string(fallback);
    };
    auto num = [&](const char *key, double fallback) {
        const Json *j = req.find(key);
        return j ? j->number() : fallback;
    };
    o.device = str("device", "gpu");
    o.method = str("method", "auto");
    o.time_limit = num("time_limit", 60.);
    o.tol = num("tol", 1e-6);
    o.threads = int(num("threads", 1));
    if (const Json *it = req.find("iterations")) o.iteration_limit = it->integer();
    const Json *ws = req.find("warm_start");
    const bool warm = ws ? ws->boolean() : true;
    if (warm) {
        o.initial_x = previous_.x;
        o.initial_y = previous_.y;
    }
    interrupted = 0;
    Result result = nyk::solve(model_, o);
    Json out = result_json(model_, result);
    Json info = Json::object();
    info["warm_start_requested"] = warm;
    info["model_retained"] = true;
    out["session"] = std::move(info);
    if (result.accuracy.finite) previous_ = std::move(result);
    return out;
}

}  // namespace nyk

struct vantage_session {
    nyk::Session engine;
    std::string output;
    std::mutex mutex;
    explicit vantage_session(const char *path) : engine(path) {}
};

namespace {
thread_local char last_error[2048] = {};
void set_error(const char *msg) { std::snprintf(last_error, sizeof last_error, "%s", msg); }
}  // namespace

extern "C" {

unsigned vantage_abi_version(void) { return 1; }

vantage_session *vantage_open(const char *path) {
    try {
        last_error[0] = '\0';
        if (!path) throw std::runtime_error("Null model path");
        return new vantage_session(path);
    } catch (const std::exception &e) {
        set_error(e.what());
        return nullptr;
    } catch (...) {
        set_error("Native session creation failed");
        return nullptr;
    }
}

void vantage_close(vantage_session *session) { delete session; }

const char *vantage_last_error(void) { return last_error; }

const char *vantage_request(vantage_session *session, const char *request) {
    try {
        last_error[0] = '\0';
        if (!session || !request) throw std::runtime_error("Null session or request");
        std::lock_guard<std::mutex> lock(session->mutex);
        session->output = session->engine.request(nyk::Json::parse(request)).dump();
        return session->output.c_str();
    } catch (const std::exception &e) {
        set_error(e.what());
        return nullptr;
    } catch (...) {
        set_error("Native request failed");
        return nullptr;
    }
}

unsigned niryukti_abi_version(void) { return vantage_abi_version(); }
niryukti_session *niryukti_open(const char *path) { return vantage_open(path); }
void niryukti_close(niryukti_session *session) { vantage_close(session); }
const char *niryukti_last_error(void) { return vantage_last_error(); }
const char *niryukti_request(niryukti_session *session, const char *request) { return vantage_request(session, request); }
}
