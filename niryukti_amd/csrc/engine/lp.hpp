// Host-side LP model, options and results of the MI355X PDHG engine.
//
// The types keep the reference's public semantics (include/vantage/vantage.hpp:
// Sparse 17-24, Model 26-39, Options 40-66, Accuracy 78-82, Result 96-123) for
// the continuous-LP PDHG path, so that results, JSON keys and defaults line up
// with Niryukti.  Quadratic and integer models are outside this engine.
#pragma once
#include <atomic>
#include <cmath>
#include <cstdint>
#include <limits>
#include <memory>
#include <string>
#include <vector>

namespace nyk {

constexpr double kInf = std::numeric_limits<double>::infinity();

struct Entry {
    int64_t row, col;
    double value;
};

// Compressed sparse rows with 64-bit offsets and column indices on the host;
// the device copy narrows to 32-bit indices when they fit.
struct Csr {
    int64_t rows = 0, cols = 0;
    std::vector<int64_t> ptr, index;
    std::vector<double> value;
    int64_t nnz() const { return int64_t(value.size()); }
    // COO -> CSR: rows ascending, columns ascending, duplicates summed in input
    // order, exact zeros dropped (src/model.cpp:11-38).
    static Csr from_entries(int64_t rows, int64_t cols, std::vector<Entry> entries);
    Csr transposed() const;                                   // src/model.cpp:39-58
    void multiply(const double *x, double *y) const;          // src/model.cpp:59-73
};

struct Model {
    std::string name = "model";
    Csr A;
    std::vector<double> c, lb, ub, rl, ru;
    std::vector<std::string> names, row_names;  // may be empty: x<j> / r<i> implied
    double offset = 0, sense = 1;               // reported objective = sense * canonical
    int64_t rows() const { return A.rows; }
    int64_t cols() const { return A.cols; }
    void validate() const;            // src/model.cpp:74-109 (LP subset)
    std::string fingerprint() const;  // src/model.cpp:274-309
    std::string column_name(int64_t j) const { return names.empty() ? "x" + std::to_string(j) : names[j]; }
    std::string row_name(int64_t i) const { return row_names.empty() ? "r" + std::to_string(i) : row_names[i]; }
};

struct Options {
    std::string device = "gpu";
    std::string method = "pdhg";           // pdhg | halpern | rhpdhg | r2hpdhg
    std::string scaling = "ruiz";          // ruiz | combined
    std::string primal_weight = "displacement";  // displacement | pid
    std::string matrix_precision = "fp64";       // fp64 | mixed (FP32 matrix values)
    std::string gpu_indices = "auto";            // auto | 32 | 64
    double tol = 1e-6, time_limit = 60;
    int64_t iteration_limit = 100000;
    int check_every = 100, scaling_passes = 5, threads = 1;
    bool presolve = true, restart = true, adaptive = true, verbose = false;
    bool graphs = true;      // hipGraph-captured chunks
    // Candidate checks: -1 on the device for models with >= 100,000 nonzeros,
    // 0 on the host every chunk, 1 on the device (the host still certifies).
    int device_checks = -1;
    // Ruiz scaling, objective scaling and the transpose: -1 on the device for
    // models with >= 100,000 nonzeros and "ruiz" scaling, 0 on the host, 1 on
    // the device (presolve always runs on the host).
    int device_prepare = -1;
    // Power iterations for an estimate of ||A||_2 (src/research.cpp:6-50);
    // 0 (the default) skips it.  With adaptive steps the initial step uses
    // min(bound, 1.05 x estimate) (src/solver.cpp:315-321).
    int power_iterations = 0;
    // Feasibility polishing (src/solver.cpp:634-690): off by default.
    bool polishing = false;
    std::vector<double> initial_x, initial_y;
    // Continuous checkpoints (src/solver.cpp:384-509): written to
    // checkpoint_path every checkpoint_nodes iterations, read by resume_path.
    std::string checkpoint_path, resume_path;
    int64_t checkpoint_nodes = 100;
    std::shared_ptr<std::atomic<bool>> cancellation;
};

struct Accuracy {
    double objective = kInf, primal = kInf, dual = kInf, gap = kInf, kkt = kInf;
    double primal_absolute = kInf, dual_absolute = kInf, complementarity = kInf, lower_bound = -kInf;
    bool finite = false;
};

struct Result {
    std::string status = "UNKNOWN", message, backend = "hip", device_name;
    std::vector<double> x, y, infeasibility_ray;
    double certificate_margin = 0;
    Accuracy accuracy;
    double seconds = 0, preprocess_seconds = 0, transfer_seconds = 0, iteration_seconds = 0,
           verification_seconds = 0;
    int64_t iterations = 0, restarts = 0, rejected_steps = 0, removed_columns = 0, removed_rows = 0;
    int64_t weight_updates = 0, host_candidate_checks = 0, device_candidate_checks = 0;
    int gpu_index_bits = 0;
    bool graph_execution = false;
    bool device_prepare = false;  // scaling and transpose ran on the device
    double operator_norm_estimate = 0;
    int64_t polishing_iterations = 0, polishing_attempts = 0;
    // Decision trace: the first step and primal weight (src/solver.cpp:299-333),
    // and the iteration of the first restart with the weight after it
    // (src/solver.cpp:692-725; 0 and 0 when no restart happened).
    double initial_step = 0, initial_weight = 0, first_restart_weight = 0;
    int64_t first_restart_iteration = 0;
    // --matrix-precision fp32: iterations run in FP32 arithmetic before the
    // FP64 refinement phase (all of them when it never started)
    int64_t fp32_iterations = 0;
    std::string matrix_precision = "fp64", method_selected, device_reason;
    double estimated_gpu_bytes = 0;
    // The request as the caller made it (reported under "selection").
    std::string method_requested, scaling = "ruiz";
    bool automatic = false;
    double tol = 1e-6, time_limit = 60;
    int64_t iteration_limit = 100000;
    int scaling_passes = 5;
};

// Presolved and scaled LP plus the maps back to the original model
// (src/internal.hpp:16-25).
struct Prepared {
    Model model;
    std::vector<int64_t> cols, rows, lower_rows, upper_rows;
    std::vector<double> fixed, column_scale, row_scale;
    double objective_scale = 1;
    std::string failure, reason;
    std::vector<double> failure_ray;
    std::vector<double> restore_x(const std::vector<double> &x) const;                 // preprocess.cpp:121-126
    std::vector<double> restore_y(const std::vector<double> &y, size_t original_rows) const;  // :127-132
};

Prepared prepare(const Model &, const Options &);

// Power-iteration estimate of ||A||_2 (src/research.cpp:6-50): 0 when
// iterations < 1 or the matrix is empty.  Same seed, start vector and
// normalisation order as the reference, so the estimate is the same.
double power_norm(const Csr &A, int iterations);
// The initial step's operator-norm bound sqrt(||A||_1 ||A||_inf), lowered to
// 1.05 x the power estimate for adaptive steps (src/solver.cpp:299-321).
double step_norm(const Model &scaled, const Options &o, double *estimate);

class Verifier {
    const Model &model_;
    Csr transpose_;
    std::vector<double> activity_, transpose_product_;

  public:
    explicit Verifier(const Model &model);
    const Csr &transpose() const { return transpose_; }
    Accuracy evaluate(const std::vector<double> &x, const std::vector<double> &y, bool safe_bound = true);
    double infeasibility_bound(const std::vector<double> &y) const;
};
Accuracy verify(const Model &, const std::vector<double> &x, const std::vector<double> &y);

void validate_options(const Options &);
Result solve(const Model &, const Options &);

}  // namespace nyk
