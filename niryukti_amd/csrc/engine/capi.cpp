// Array-level C-ABI of the engine (include/niryukti_engine.h).
#include "niryukti_engine.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>

#include "instances.hpp"
#include "io.hpp"
#include "lp.hpp"

namespace {
thread_local std::string g_error;

nyk::Model to_model(const niryukti_lp &lp) {
    if (lp.rows < 0 || lp.cols < 0 || !lp.ptr) throw std::runtime_error("Invalid LP arrays");
    nyk::Model m;
    const int64_t nnz = lp.ptr[lp.rows];
    if (nnz < 0 || (nnz > 0 && (!lp.index || !lp.value))) throw std::runtime_error("Invalid CSR arrays");
    if ((lp.cols > 0 && (!lp.c || !lp.lb || !lp.ub)) || (lp.rows > 0 && (!lp.rl || !lp.ru)))
        throw std::runtime_error("Missing LP vectors");
    m.A.rows = lp.rows;
    m.A.cols = lp.cols;
    m.A.ptr.assign(lp.ptr, lp.ptr + lp.rows + 1);
    m.A.index.assign(lp.index, lp.index + nnz);
    m.A.value.assign(lp.value, lp.value + nnz);
    m.c.assign(lp.c, lp.c + lp.cols);
    m.lb.assign(lp.lb, lp.lb + lp.cols);
    m.ub.assign(lp.ub, lp.ub + lp.cols);
    m.rl.assign(lp.rl, lp.rl + lp.rows);
    m.ru.assign(lp.ru, lp.ru + lp.rows);
    m.offset = lp.offset;
    m.sense = lp.sense == 0 ? 1 : lp.sense;
    return m;
}

nyk::Options to_options(const niryukti_options *in, const niryukti_lp &lp) {
    nyk::Options o;
    if (!in) return o;
    auto str = [](const char *s, const char *d) { return std::string(s ? s : d); };
    o.device = str(in->device, "gpu");
    o.method = str(in->method, "pdhg");
    o.scaling = str(in->scaling, "ruiz");
    o.primal_weight = str(in->primal_weight, "displacement");
    o.matrix_precision = str(in->matrix_precision, "fp64");
    o.gpu_indices = str(in->gpu_indices, "auto");
    o.tol = in->tol;
    o.time_limit = in->time_limit;
    o.iteration_limit = in->iteration_limit;
    o.check_every = in->check_every;
    o.scaling_passes = in->scaling_passes;
    o.threads = in->threads;
    o.presolve = in->presolve != 0;
    o.restart = in->restart != 0;
    o.adaptive = in->adaptive != 0;
    o.graphs = in->graphs != 0;
    o.verbose = in->verbose != 0;
    o.device_checks = in->device_checks < 0 ? -1 : in->device_checks > 0 ? 1 : 0;
    o.device_prepare = in->device_prepare < 0 ? -1 : in->device_prepare > 0 ? 1 : 0;
    o.power_iterations = in->power_iterations;
    o.polishing = in->polishing != 0;
    if (in->initial_x) o.initial_x.assign(in->initial_x, in->initial_x + lp.cols);
    if (in->initial_y) o.initial_y.assign(in->initial_y, in->initial_y + lp.rows);
    if (in->checkpoint_path) o.checkpoint_path = in->checkpoint_path;
    if (in->resume_path) o.resume_path = in->resume_path;
    o.checkpoint_nodes = in->checkpoint_nodes;
    return o;
}
}  // namespace

struct niryukti_model {
    nyk::Model m;
    std::string fingerprint;
};

struct niryukti_instance {
    nyk_gen::Instance p;
};

extern "C" {

void niryukti_default_options(niryukti_options *o) {
    std::memset(o, 0, sizeof *o);
    o->device = "gpu";
    o->method = "pdhg";
    o->scaling = "ruiz";
    o->primal_weight = "displacement";
    o->matrix_precision = "fp64";
    o->gpu_indices = "auto";
    o->tol = 1e-6;
    o->time_limit = 60;
    o->iteration_limit = 100000;
    o->check_every = 100;
    o->scaling_passes = 5;
    o->threads = 1;
    o->presolve = 1;
    o->restart = 1;
    o->adaptive = 1;
    o->graphs = 1;
    o->device_checks = -1;
    o->device_prepare = -1;
    o->power_iterations = 0;
    o->polishing = 0;
    o->checkpoint_path = nullptr;
    o->resume_path = nullptr;
    o->checkpoint_nodes = 100;
}

int niryukti_prepare_lp(const niryukti_lp *lp, const niryukti_options *opts, int64_t *out_rows, int64_t *out_cols,
                        int64_t *ptr, int64_t *index, double *value, double *c, double *lb, double *ub, double *rl,
                        double *ru, double *column_scale, double *row_scale, int64_t *col_map, int64_t *row_map,
                        double *objective_scale, double *offset) {
    try {
        g_error.clear();
        if (!lp || !out_rows || !out_cols || !objective_scale || !offset) throw std::runtime_error("Null argument");
        const nyk::Model m = to_model(*lp);
        const nyk::Options o = to_options(opts, *lp);
        const nyk::Prepared p = nyk::prepare(m, o);
        if (!p.failure.empty()) {
            g_error = p.failure + ": " + p.reason;
            return 1;
        }
        const nyk::Model &q = p.model;
        *out_rows = q.rows();
        *out_cols = q.cols();
        auto put = [](auto *dst, const auto &v) {
            if (dst && !v.empty()) std::copy(v.begin(), v.end(), dst);
        };
        put(ptr, q.A.ptr);
        put(index, q.A.index);
        put(value, q.A.value);
        put(c, q.c);
        put(lb, q.lb);
        put(ub, q.ub);
        put(rl, q.rl);
        put(ru, q.ru);
        put(column_scale, p.column_scale);
        put(row_scale, p.row_scale);
        put(col_map, p.cols);
        put(row_map, p.rows);
        *objective_scale = p.objective_scale;
        *offset = q.offset;
        return 0;
    } catch (const std::exception &e) {
        g_error = e.what();
        return -1;
    }
}

int niryukti_verify_lp(const niryukti_lp *lp, const double *x, const double *y, int safe_bound,
                       niryukti_accuracy *out) {
    try {
        g_error.clear();
        if (!lp || !x || !y || !out) throw std::runtime_error("Null argument");
        auto m = to_model(*lp);
        nyk::Verifier v(m);
        const auto a = v.evaluate(std::vector<double>(x, x + lp->cols), std::vector<double>(y, y + lp->rows),
                                  safe_bound != 0);
        out->objective = a.finite ? m.sense * a.objective : NAN;
        out->primal = a.primal;
        out->dual = a.dual;
        out->gap = a.gap;
        out->kkt = a.kkt;
        out->complementarity = a.complementarity;
        out->lower_bound = a.lower_bound;
        out->primal_absolute = a.primal_absolute;
        out->dual_absolute = a.dual_absolute;
        out->finite = a.finite ? 1 : 0;
        return 0;
    } catch (const std::exception &e) {
        g_error = e.what();
        return -1;
    }
}

double niryukti_power_norm(const niryukti_lp *lp, int iterations) {
    try {
        g_error.clear();
        if (!lp) throw std::runtime_error("Null argument");
        return nyk::power_norm(to_model(*lp).A, iterations);
    } catch (const std::exception &e) {
        g_error = e.what();
        return 0;
    }
}

int niryukti_solve_lp(const niryukti_lp *lp, const niryukti_options *opts, niryukti_result *out) {
    try {
        g_error.clear();
        if (!lp || !out) throw std::runtime_error("Null argument");
        auto m = to_model(*lp);
        auto o = to_options(opts, *lp);
        auto r = nyk::solve(m, o);
        std::snprintf(out->status, sizeof out->status, "%s", r.status.c_str());
        std::snprintf(out->message, sizeof out->message, "%s", r.message.c_str());
        const auto &a = r.accuracy;
        out->objective = a.finite ? m.sense * a.objective : NAN;
        out->primal_residual = a.primal;
        out->dual_residual = a.dual;
        out->relative_gap = a.gap;
        out->kkt_error = a.kkt;
        out->complementarity = a.complementarity;
        out->dual_bound = a.lower_bound;
        out->primal_absolute = a.primal_absolute;
        out->dual_absolute = a.dual_absolute;
        out->finite = a.finite ? 1 : 0;
        out->iterations = r.iterations;
        out->restarts = r.restarts;
        out->rejected_steps = r.rejected_steps;
        out->weight_updates = r.weight_updates;
        out->host_candidate_checks = r.host_candidate_checks;
        out->device_candidate_checks = r.device_candidate_checks;
        out->removed_rows = r.removed_rows;
        out->removed_columns = r.removed_columns;
        out->seconds = r.seconds;
        out->preprocess_seconds = r.preprocess_seconds;
        out->transfer_seconds = r.transfer_seconds;
        out->iteration_seconds = r.iteration_seconds;
        out->verification_seconds = r.verification_seconds;
        out->index_bits = r.gpu_index_bits;
        out->device_prepare = r.device_prepare ? 1 : 0;
        out->operator_norm_estimate = r.operator_norm_estimate;
        out->polishing_iterations = r.polishing_iterations;
        out->polishing_attempts = r.polishing_attempts;
        out->initial_step = r.initial_step;
        out->initial_weight = r.initial_weight;
        out->first_restart_weight = r.first_restart_weight;
        out->first_restart_iteration = r.first_restart_iteration;
        out->fp32_iterations = r.fp32_iterations;
        if (out->x && r.x.size() == size_t(lp->cols)) std::memcpy(out->x, r.x.data(), r.x.size() * sizeof(double));
        if (out->y && r.y.size() == size_t(lp->rows)) std::memcpy(out->y, r.y.data(), r.y.size() * sizeof(double));
        return 0;
    } catch (const std::exception &e) {
        g_error = e.what();
        return -1;
    }
}

const char *niryukti_engine_last_error(void) { return g_error.c_str(); }

niryukti_model *niryukti_model_load(const char *path, const char *json_text) {
    try {
        g_error.clear();
        if ((path == nullptr) == (json_text == nullptr)) throw std::runtime_error("Give exactly one of path and JSON text");
        auto *h = new niryukti_model;
        try {
            h->m = path ? nyk::read_model(path) : nyk::read_json_model(json_text);
            h->m.validate();
            h->fingerprint = h->m.fingerprint();
        } catch (...) {
            delete h;
            throw;
        }
        return h;
    } catch (const std::exception &e) {
        g_error = e.what();
        return nullptr;
    }
}

void niryukti_model_destroy(niryukti_model *m) { delete m; }

int niryukti_model_write(const niryukti_model *m, const char *path) {
    try {
        g_error.clear();
        if (!m || !path) throw std::runtime_error("model and path are required");
        nyk::write_model(m->m, path);
        return 0;
    } catch (const std::exception &e) {
        g_error = e.what();
        return -1;
    }
}

void niryukti_model_view(const niryukti_model *h, niryukti_lp *lp, const char **name, const char **fingerprint) {
    const nyk::Model &m = h->m;
    lp->rows = m.rows();
    lp->cols = m.cols();
    lp->ptr = m.A.ptr.data();
    lp->index = m.A.index.data();
    lp->value = m.A.value.data();
    lp->c = m.c.data();
    lp->lb = m.lb.data();
    lp->ub = m.ub.data();
    lp->rl = m.rl.data();
    lp->ru = m.ru.data();
    lp->offset = m.offset;
    lp->sense = m.sense;
    if (name) *name = m.name.c_str();
    if (fingerprint) *fingerprint = h->fingerprint.c_str();
}

const char *niryukti_result_json(const niryukti_model *h, const niryukti_result *in, const char *backend) {
    static thread_local std::string out;
    try {
        g_error.clear();
        if (!h || !in) throw std::runtime_error("Null argument");
        const nyk::Model &m = h->m;
        nyk::Result r;
        r.status = in->status;
        r.message = in->message;
        r.backend = backend ? backend : "hip";
        r.method_selected = r.method_requested = "pdhg";
        r.accuracy.finite = in->finite != 0;
        r.accuracy.objective = in->finite ? in->objective / m.sense : nyk::kInf;
        r.accuracy.primal = in->primal_residual;
        r.accuracy.dual = in->dual_residual;
        r.accuracy.gap = in->relative_gap;
        r.accuracy.kkt = in->kkt_error;
        r.accuracy.complementarity = in->complementarity;
        r.accuracy.lower_bound = in->dual_bound;
        r.accuracy.primal_absolute = in->primal_absolute;
        r.accuracy.dual_absolute = in->dual_absolute;
        r.iterations = in->iterations;
        r.restarts = in->restarts;
        r.rejected_steps = in->rejected_steps;
        r.weight_updates = in->weight_updates;
        r.host_candidate_checks = in->host_candidate_checks;
        r.device_candidate_checks = in->device_candidate_checks;
        r.removed_rows = in->removed_rows;
        r.removed_columns = in->removed_columns;
        r.seconds = in->seconds;
        r.preprocess_seconds = in->preprocess_seconds;
        r.transfer_seconds = in->transfer_seconds;
        r.iteration_seconds = in->iteration_seconds;
        r.verification_seconds = in->verification_seconds;
        r.gpu_index_bits = in->index_bits;
        r.device_prepare = in->device_prepare != 0;
        r.operator_norm_estimate = in->operator_norm_estimate;
        r.polishing_iterations = in->polishing_iterations;
        r.polishing_attempts = in->polishing_attempts;
        r.estimated_gpu_bytes = 48. * m.A.nnz() + 240. * (m.rows() + m.cols() + 2);
        if (in->x) r.x.assign(in->x, in->x + m.cols());
        if (in->y) r.y.assign(in->y, in->y + m.rows());
        out = nyk::result_json(m, r).dump();
        return out.c_str();
    } catch (const std::exception &e) {
        g_error = e.what();
        return nullptr;
    }
}

niryukti_instance *niryukti_instance_create(int config, uint64_t seed, int64_t scale, int64_t index) {
    try {
        g_error.clear();
        auto *h = new niryukti_instance;
        if (config == 2)
            h->p = nyk_gen::batch_instance(nyk_gen::batch_shapes(int(index) + 1, seed).back());
        else if (config == 3)
            h->p = nyk_gen::batch_instance(nyk_gen::serve_shape(index, seed));
        else
            h->p = nyk_gen::workload(config, seed, scale);
        return h;
    } catch (const std::exception &e) {
        g_error = e.what();
        return nullptr;
    }
}

void niryukti_instance_destroy(niryukti_instance *p) { delete p; }

void niryukti_instance_view(const niryukti_instance *h, niryukti_lp *lp, double *planted_objective,
                            const double **planted_x, const double **planted_y) {
    const auto &p = h->p;
    lp->rows = p.rows;
    lp->cols = p.cols;
    lp->ptr = p.ptr.data();
    lp->index = p.index.data();
    lp->value = p.value.data();
    lp->c = p.c.data();
    lp->lb = p.lb.data();
    lp->ub = p.ub.data();
    lp->rl = p.rl.data();
    lp->ru = p.ru.data();
    lp->offset = 0;
    lp->sense = 1;
    if (planted_objective) *planted_objective = p.objective;
    if (planted_x) *planted_x = p.x0.data();
    if (planted_y) *planted_y = p.y_opt.data();
}
}
