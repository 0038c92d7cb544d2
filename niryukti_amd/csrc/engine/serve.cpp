// Serving loop of the MI355X LP engine (include/niryukti_serve.h).
//
// One scheduler thread owns the batched engine: it moves newly submitted
// small LPs into it, steps it, and completes requests as their LPs finish.
// Grid-route requests go to a pool of worker threads that each run one
// single-LP solve at a time, admitted against the grid share of the HBM
// budget.  Waiters block on a condition variable per server.
#include "niryukti_serve.h"

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "lp.hpp"
#include "niryukti_batch.h"

namespace {

thread_local std::string g_error;
using Clock = std::chrono::steady_clock;

struct Request {
    int64_t id = 0;
    int route = 0;  // 0 batched, 1 grid
    int64_t rows = 0, cols = 0;
    std::vector<int64_t> ptr, index;
    std::vector<double> value, c, lb, ub, rl, ru, x0, y0;
    double offset = 0, sense = 1;
    bool own_options = false;
    niryukti_options options{};
    std::string option_text[8];  // storage for the option strings
    Clock::time_point submitted, completed;
    bool done = false;
    std::string error;
    niryukti_result result{};
    std::vector<double> x, y;
    int64_t ticket = -1;
    int64_t grid_bytes = 0;

    niryukti_lp view() const {
        niryukti_lp lp{};
        lp.rows = rows;
        lp.cols = cols;
        lp.ptr = ptr.data();
        lp.index = index.data();
        lp.value = value.data();
        lp.c = c.data();
        lp.lb = lb.data();
        lp.ub = ub.data();
        lp.rl = rl.data();
        lp.ru = ru.data();
        lp.offset = offset;
        lp.sense = sense;
        return lp;
    }
};

// Copies an options struct so that its strings and warm-start pointers stay
// valid inside the request.
void keep_options(Request &q, const niryukti_options &o) {
    q.options = o;
    const char *const *src[] = {&o.device, &o.method, &o.scaling, &o.primal_weight, &o.matrix_precision,
                                &o.gpu_indices, &o.checkpoint_path, &o.resume_path};
    const char **dst[] = {&q.options.device,           &q.options.method,      &q.options.scaling,
                          &q.options.primal_weight,    &q.options.matrix_precision,
                          &q.options.gpu_indices,      &q.options.checkpoint_path,
                          &q.options.resume_path};
    for (int k = 0; k < 8; ++k) {
        if (*src[k]) {
            q.option_text[k] = *src[k];
            *dst[k] = q.option_text[k].c_str();
        } else {
            *dst[k] = nullptr;
        }
    }
    q.options.initial_x = nullptr;  // warm starts travel as x0 / y0
    q.options.initial_y = nullptr;
}

class Server {
  public:
    Server(const niryukti_options *opts, const niryukti_server_limits &l) : limits_(l) {
        if (!(l.batch_fraction > 0 && l.batch_fraction < 1) || l.hbm_budget <= 0 || l.grid_workers < 1 ||
            l.grid_workers > 16 || l.checks_per_launch < 1 || l.max_slots < 1 || l.large_nonzeros < 0 ||
            l.batch_max_size < 0)
            throw std::runtime_error("Invalid server limits");
        niryukti_default_options(&defaults_);
        if (opts) defaults_ = *opts;
        keep_options(default_store_, defaults_);
        defaults_ = default_store_.options;
        grid_budget_ = int64_t(double(l.hbm_budget) * (1 - l.batch_fraction));
        scheduler_ = std::thread([this] { schedule(); });
        for (int w = 0; w < l.grid_workers; ++w) workers_.emplace_back([this] { grid_loop(); });
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> lock(mu_);
            stopping_ = true;
        }
        work_cv_.notify_all();
        grid_cv_.notify_all();
        scheduler_.join();
        for (auto &w : workers_) w.join();
        if (batch_) niryukti_batch_destroy(batch_);
    }

    int64_t submit(const niryukti_lp &lp, const double *x0, const double *y0, int route, const niryukti_options *opts) {
        if (lp.rows < 0 || lp.cols < 0 || !lp.ptr) throw std::runtime_error("Invalid LP arrays");
        auto q = std::make_unique<Request>();
        q->rows = lp.rows;
        q->cols = lp.cols;
        const int64_t nnz = lp.ptr[lp.rows];
        if (nnz < 0 || (nnz && (!lp.index || !lp.value)) || (lp.cols && (!lp.c || !lp.lb || !lp.ub)) ||
            (lp.rows && (!lp.rl || !lp.ru)))
            throw std::runtime_error("Invalid LP arrays");
        q->ptr.assign(lp.ptr, lp.ptr + lp.rows + 1);
        q->index.assign(lp.index, lp.index + nnz);
        q->value.assign(lp.value, lp.value + nnz);
        q->c.assign(lp.c, lp.c + lp.cols);
        q->lb.assign(lp.lb, lp.lb + lp.cols);
        q->ub.assign(lp.ub, lp.ub + lp.cols);
        q->rl.assign(lp.rl, lp.rl + lp.rows);
        q->ru.assign(lp.ru, lp.ru + lp.rows);
        q->offset = lp.offset;
        q->sense = lp.sense == 0 ? 1 : lp.sense;
        // the checks niryukti_solve_lp / niryukti_batch_submit make, here, so
        // that a bad request is refused at once and never reaches a route
        {
            nyk::Model m;
            m.A.rows = q->rows;
            m.A.cols = q->cols;
            m.A.ptr = q->ptr;
            m.A.index = q->index;
            m.A.value = q->value;
            m.c = q->c;
            m.lb = q->lb;
            m.ub = q->ub;
            m.rl = q->rl;
            m.ru = q->ru;
            m.validate();
        }
        if (x0) q->x0.assign(x0, x0 + lp.cols);
        if (y0) q->y0.assign(y0, y0 + lp.rows);
        for (double v : q->x0)
            if (!std::isfinite(v)) throw std::runtime_error("Nonfinite warm start");
        for (double v : q->y0)
            if (!std::isfinite(v)) throw std::runtime_error("Nonfinite warm start");
        if (route < -1 || route > 1) throw std::runtime_error("Route must be -1 (by size), 0 (batched) or 1 (grid)");
        if (opts) {
            keep_options(*q, *opts);
            q->own_options = true;
        }
        const bool large = nnz >= limits_.large_nonzeros || lp.rows + lp.cols > limits_.batch_max_size;
        q->route = route >= 0 ? route : (large || opts ? 1 : 0);
        if (q->route == 0 && opts) throw std::runtime_error("Requests with options of their own take the grid route");
        q->grid_bytes = int64_t(2 * (48.0 * double(nnz) + 240.0 * double(lp.rows + lp.cols + 2)));
        q->x.assign(size_t(lp.cols), 0.0);
        q->y.assign(size_t(lp.rows), 0.0);
        q->submitted = Clock::now();
        std::lock_guard<std::mutex> lock(mu_);
        if (stopping_) throw std::runtime_error("Server is closing");
        q->id = next_id_++;
        Request *raw = q.get();
        requests_[raw->id] = std::move(q);
        ++submitted_;
        if (raw->route == 0) {
            ++batch_routed_;
            incoming_.push_back(raw);
            work_cv_.notify_one();
        } else {
            ++grid_routed_;
            grid_queue_.push_back(raw);
            grid_cv_.notify_one();
        }
        return raw->id;
    }

    int wait(int64_t id, double timeout, niryukti_result *out, double *latency, int *route) {
        std::unique_lock<std::mutex> lock(mu_);
        auto it = requests_.find(id);
        if (it == requests_.end()) throw std::runtime_error("Unknown request");
        Request &q = *it->second;
        auto ready = [&] { return q.done; };
        if (timeout < 0)
            done_cv_.wait(lock, ready);
        else if (!done_cv_.wait_for(lock, std::chrono::duration<double>(timeout), ready))
            return 0;
        if (route) *route = q.route;
        if (latency) *latency = std::chrono::duration<double>(q.completed - q.submitted).count();
        if (!q.error.empty()) throw std::runtime_error(q.error);
        if (out) {
            double *x = out->x, *y = out->y;
            *out = q.result;
            out->x = x;
            out->y = y;
            if (x && !q.x.empty()) std::memcpy(x, q.x.data(), q.x.size() * sizeof(double));
            if (y && !q.y.empty()) std::memcpy(y, q.y.data(), q.y.size() * sizeof(double));
        }
        return 1;
    }

    void release(int64_t id) {
        std::lock_guard<std::mutex> lock(mu_);
        auto it = requests_.find(id);
        if (it == requests_.end()) throw std::runtime_error("Unknown request");
        if (!it->second->done) throw std::runtime_error("Request has not finished");
        requests_.erase(it);
    }

    void stats(niryukti_server_stats *s) {
        std::lock_guard<std::mutex> lock(mu_);
        s->submitted = submitted_;
        s->finished = finished_;
        s->batch_routed = batch_routed_;
        s->grid_routed = grid_routed_;
        s->grid_running = grid_running_;
        s->grid_bytes_in_use = grid_used_;
        s->grid_budget_bytes = grid_budget_;
        s->batch_rounds = batch_rounds_;
    }

  private:
    void complete(Request &q, const std::string &error) {
        {
            std::lock_guard<std::mutex> lock(mu_);
            q.error = error;
            q.completed = Clock::now();
            q.done = true;
            ++finished_;
        }
        done_cv_.notify_all();
    }

    // The scheduler: batched route only.
    void schedule() {
        std::map<int64_t, Request *> tickets;  // batch ticket -> request
        std::vector<int64_t> finished(4096);
        for (;;) {
            std::vector<Request *> take;
            {
                std::unique_lock<std::mutex> lock(mu_);
                work_cv_.wait(lock, [&] { return stopping_ || !incoming_.empty() || !tickets.empty(); });
                if (stopping_ && incoming_.empty() && tickets.empty()) break;
                take.assign(incoming_.begin(), incoming_.end());
                incoming_.clear();
            }
            for (Request *q : take) {
                if (!batch_ && !open_batch(*q)) continue;
                const niryukti_lp lp = q->view();
                int64_t ticket = -1;
                if (niryukti_batch_submit(batch_, &lp, q->x0.empty() ? nullptr : q->x0.data(),
                                          q->y0.empty() ? nullptr : q->y0.data(), &ticket) != 0) {
                    complete(*q, niryukti_batch_last_error());
                    continue;
                }
                q->ticket = ticket;
                tickets[ticket] = q;
            }
            if (tickets.empty()) continue;
            int64_t running = 0;
            if (niryukti_batch_step(batch_, limits_.checks_per_launch, &running) != 0) {
                // the engine itself failed: every request it holds fails with it
                const std::string error = std::string("Batched engine failed: ") + niryukti_batch_last_error();
                for (auto &kv : tickets) complete(*kv.second, error);
                tickets.clear();
                niryukti_batch_destroy(batch_);
                batch_ = nullptr;
                continue;
            }
            {
                std::lock_guard<std::mutex> lock(mu_);
                ++batch_rounds_;
            }
            for (;;) {
                int64_t count = 0;
                if (niryukti_batch_finished(batch_, finished.data(), int64_t(finished.size()), &count) != 0 || !count)
                    break;
                for (int64_t k = 0; k < count; ++k) {
                    auto it = tickets.find(finished[size_t(k)]);
                    if (it == tickets.end()) continue;
                    Request &q = *it->second;
                    q.result = niryukti_result{};
                    q.result.x = q.x.data();
                    q.result.y = q.y.data();
                    std::string error;
                    if (niryukti_batch_result(batch_, q.ticket, &q.result) != 1) error = niryukti_batch_last_error();
                    q.result.x = nullptr;
                    q.result.y = nullptr;
                    (void)niryukti_batch_release(batch_, q.ticket);
                    tickets.erase(it);
                    complete(q, error.empty() ? std::string() : "Batched result: " + error);
                }
            }
        }
    }

    // The batched engine is created with the first batched request (the
    // server itself needs no device until then).
    bool open_batch(Request &q) {
        const int64_t arena = int64_t(double(limits_.hbm_budget) * limits_.batch_fraction);
        if (niryukti_batch_create(&defaults_, arena, limits_.max_slots, &batch_) != 0) {
            batch_ = nullptr;
            complete(q, niryukti_batch_last_error());
            return false;
        }
        return true;
    }

    void grid_loop() {
        for (;;) {
            Request *q = nullptr;
            {
                std::unique_lock<std::mutex> lock(mu_);
                grid_cv_.wait(lock, [&] { return stopping_ || !grid_queue_.empty(); });
                if (grid_queue_.empty()) return;  // stopping, and nothing left to run
                q = grid_queue_.front();
                grid_queue_.pop_front();
                // admit when it fits, or alone when it never could
                grid_cv_.wait(lock, [&] { return grid_running_ == 0 || grid_used_ + q->grid_bytes <= grid_budget_; });
                grid_used_ += q->grid_bytes;
                ++grid_running_;
            }
            niryukti_options o = q->own_options ? q->options : defaults_;
            o.initial_x = q->x0.empty() ? nullptr : q->x0.data();
            o.initial_y = q->y0.empty() ? nullptr : q->y0.data();
            const niryukti_lp lp = q->view();
            q->result = niryukti_result{};
            q->result.x = q->x.data();
            q->result.y = q->y.data();
            std::string error;
            if (niryukti_solve_lp(&lp, &o, &q->result) != 0) error = niryukti_engine_last_error();
            q->result.x = nullptr;
            q->result.y = nullptr;
            {
                std::lock_guard<std::mutex> lock(mu_);
                grid_used_ -= q->grid_bytes;
                --grid_running_;
            }
            grid_cv_.notify_all();
            complete(*q, error);
        }
    }

    niryukti_server_limits limits_;
    niryukti_options defaults_{};
    Request default_store_;  // keeps the default option strings alive
    niryukti_batch *batch_ = nullptr;
    std::mutex mu_;
    std::condition_variable work_cv_, grid_cv_, done_cv_;
    std::deque<Request *> incoming_, grid_queue_;
    std::map<int64_t, std::unique_ptr<Request>> requests_;
    int64_t next_id_ = 1, submitted_ = 0, finished_ = 0, batch_routed_ = 0, grid_routed_ = 0;
    int64_t grid_running_ = 0, grid_used_ = 0, grid_budget_ = 0, batch_rounds_ = 0;
    bool stopping_ = false;
    std::thread scheduler_;
    std::vector<std::thread> workers_;
};

template <class F>
int guarded(F &&f) {
    try {
        g_error.clear();
        return f();
    } catch (const std::exception &e) {
        g_error = e.what();
        return -1;
    }
}

}  // namespace

struct niryukti_server {
    Server impl;
    niryukti_server(const niryukti_options *o, const niryukti_server_limits &l) : impl(o, l) {}
};

extern "C" {

void niryukti_server_default_limits(niryukti_server_limits *l) {
    l->hbm_budget = int64_t(64) << 30;
    l->batch_fraction = 0.75;
    l->large_nonzeros = 200000;
    l->grid_workers = 2;
    l->checks_per_launch = 1;
    l->max_slots = 4096;
    l->batch_max_size = 8000;
}

int niryukti_server_create(const niryukti_options *opts, const niryukti_server_limits *limits, niryukti_server **out) {
    return guarded([&] {
        if (!out) throw std::runtime_error("Null argument");
        niryukti_server_limits l;
        niryukti_server_default_limits(&l);
        if (limits) l = *limits;
        *out = new niryukti_server(opts, l);
        return 0;
    });
}

void niryukti_server_destroy(niryukti_server *s) { delete s; }

int niryukti_server_submit(niryukti_server *s, const niryukti_lp *lp, const double *x0, const double *y0, int route,
                           const niryukti_options *opts, int64_t *request) {
    return guarded([&] {
        if (!s || !lp || !request) throw std::runtime_error("Null argument");
        *request = s->impl.submit(*lp, x0, y0, route, opts);
        return 0;
    });
}

int niryukti_server_wait(niryukti_server *s, int64_t request, double timeout_seconds, niryukti_result *r,
                         double *latency, int *route) {
    return guarded([&] {
        if (!s) throw std::runtime_error("Null argument");
        return s->impl.wait(request, timeout_seconds, r, latency, route);
    });
}

int niryukti_server_release(niryukti_server *s, int64_t request) {
    return guarded([&] {
        if (!s) throw std::runtime_error("Null argument");
        s->impl.release(request);
        return 0;
    });
}

int niryukti_server_stats_get(niryukti_server *s, niryukti_server_stats *out) {
    return guarded([&] {
        if (!s || !out) throw std::runtime_error("Null argument");
        s->impl.stats(out);
        return 0;
    });
}

const char *niryukti_server_last_error(void) { return g_error.c_str(); }
}
