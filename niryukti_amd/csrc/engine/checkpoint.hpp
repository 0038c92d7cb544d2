// Continuous-solve checkpoints (the reference's `vantage-continuous-1` files,
// src/solver.cpp:384-509): the driver's loop state at the top of a chunk,
// written atomically every `checkpoint_nodes` iterations and read back by
// --resume, after which the solve continues exactly as it would have.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "json.hpp"
#include "lp.hpp"

namespace nyk {

struct LoopState {
    std::vector<std::vector<double>> backend;  // npdhg_snapshot, split into its arrays
    std::vector<double> best_x, best_y;        // best pair so far, original space
    std::vector<double> epoch_x, epoch_y;      // restart epoch, scaled space (host checks)
    std::vector<double> pid;                   // PID controller: integral, previous error, initialised
    double step = 0, weight = 0, restart_kkt = 0;
    int64_t since_restart = 0, next_polish = 100, previous_rejected = 0;
    int64_t iterations = 0, restarts = 0, rejected_steps = 0, weight_updates = 0;
    int64_t host_checks = 0, device_checks = 0, polishing_iterations = 0, polishing_attempts = 0;
    int64_t first_restart_iteration = 0;
    double first_restart_weight = 0;
    bool best_certified = true;  // host checks: the best pair's figures include the certified bound
    double best_kkt = 0;         // device checks: rel-KKT of the best candidate on the device
    bool have_best = false;      // device checks: the device holds a best candidate
    bool fp32_phase = false;     // fp32 precision: still in the FP32 arithmetic phase
    double fp32_best = kInf;     // fp32 precision: best rel-KKT seen in that phase
    int64_t fp32_since = 0, fp32_iterations = 0;
};

// The solver settings a checkpoint belongs to (a resumed solve must use the
// same ones): the reference's keys plus whether checks run on the device.
Json checkpoint_configuration(const Options &o, bool device_checks);

// Writes to `path`.tmp, then renames over `path`.
void write_checkpoint(const std::string &path, const Model &original, const Json &configuration, const LoopState &s);

// Reads a checkpoint written for `original` with `configuration`; throws
// when the schema, model fingerprint or configuration differ, or a value is
// missing, NaN or out of range.  Array lengths are checked by the caller.
LoopState read_checkpoint(const std::string &path, const Model &original, const Json &configuration);

// The schema, fingerprint and every configuration entry except where the
// checks run (before any device is touched).
void check_checkpoint_model(const std::string &path, const Model &original, const Json &configuration);

}  // namespace nyk
