// Model files, result JSON and solution files (the Niryukti drop-in formats).
#pragma once
#include <string>

#include "json.hpp"
#include "lp.hpp"

namespace nyk {

// .json (native format, src/io.cpp:84-140) or .mps (src/io.cpp:299-585)
// by extension; LP only (integer variables and quadratic objectives are
// rejected).  Returns the canonical minimisation model.
Model read_model(const std::string &path);
Model read_json_model(const std::string &text);
Model read_mps_model(const std::string &text);

// Writes .json (native format, src/io.cpp:765-798) or .mps (src/io.cpp:815-901)
// by extension: the reference's `convert INPUT OUTPUT`.
void write_model(const Model &, const std::string &path);

// Result document with the reference's keys (src/io.cpp:902-995).
Json result_json(const Model &, const Result &);

struct Solution {
    std::vector<double> x, y, infeasibility_ray;
    std::string status;
    double objective = std::numeric_limits<double>::quiet_NaN();  // reported (sense-restored)
};
// Reads a result document for warm starts and verification
// (src/io.cpp:1002-1032): fingerprint must match unless allow_model_change,
// in which case names and ordering must match.
Solution read_solution(const Model &, const std::string &path, bool allow_model_change = false);

std::string engine_version();

}  // namespace nyk
