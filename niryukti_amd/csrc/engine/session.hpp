// A model held in memory between requests: bound/objective updates and
// warm-started solves (Niryukti's session protocol).
#pragma once
#include <string>

#include "io.hpp"
#include "json.hpp"
#include "lp.hpp"

namespace nyk {

class Session {
  public:
    explicit Session(const std::string &path);
    std::string fingerprint() const { return model_.fingerprint(); }
    const Model &model() const { return model_; }
    // One JSON request -> one JSON response; throws std::runtime_error on a
    // malformed or rejected request.
    Json request(const Json &req);

  private:
    Json update(const Json &req);
    Json solve(const Json &req);
    Model model_;
    Result previous_;
};

}  // namespace nyk
