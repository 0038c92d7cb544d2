// Host-code sanitizer harness (built by tests/sanitize/Makefile with
// -fsanitize=address,undefined or -fsanitize=thread, no GPU): drives the
// engine's host paths through their C-ABIs and internal interfaces and exits
// non-zero on any wrong answer; the sanitizers abort on any memory error,
// undefined behaviour or data race they find.
//
//   harness NETLIB_DIR TMP_DIR
//
// Covered: model readers and writer (MPS, native JSON), fingerprints, the host
// prepare (presolve + scaling) on the seeded workloads, the certifying
// verifier, the session protocol, checkpoint files, the serving loop's
// scheduler / grid workers / waiters under concurrent submitters, and the
// certify-beside-the-next-round handoff the batched engine uses.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "engine/checkpoint.hpp"
#include "engine/lp.hpp"
#include "engine/parallel.hpp"
#include "niryukti/c_api.h"
#include "niryukti_engine.h"
#include "niryukti_serve.h"

namespace {
int failures = 0;
void expect(bool ok, const char *what) {
    if (!ok) {
        std::fprintf(stderr, "FAIL: %s\n", what);
        ++failures;
    }
}

void models(const std::string &netlib, const std::string &tmp) {
    for (const char *name : {"afiro", "adlittle", "israel"}) {
        const std::string path = netlib + "/" + name + ".mps";
        niryukti_model *m = niryukti_model_load(path.c_str(), nullptr);
        expect(m != nullptr, "MPS model loads");
        if (!m) continue;
        niryukti_lp lp{};
        const char *model_name = nullptr, *fp = nullptr;
        niryukti_model_view(m, &lp, &model_name, &fp);
        expect(lp.rows > 0 && lp.cols > 0 && fp && std::strlen(fp) > 0, "model view");
        const std::string out = tmp + "/" + name + ".json";
        expect(niryukti_model_write(m, out.c_str()) == 0, "model write");
        niryukti_model *again = niryukti_model_load(out.c_str(), nullptr);
        expect(again != nullptr, "written model reloads");
        if (again) {
            niryukti_lp lp2{};
            const char *n2 = nullptr, *fp2 = nullptr;
            niryukti_model_view(again, &lp2, &n2, &fp2);
            expect(std::string(fp) == fp2, "JSON round trip keeps the fingerprint");
            niryukti_model_destroy(again);
        }
        // prepare and verify the start point
        std::vector<int64_t> ptr(size_t(lp.rows + 1)), index(size_t(lp.ptr[lp.rows])), cmap(size_t(lp.cols)),
            rmap(size_t(lp.rows));
        std::vector<double> value(index.size()), c(size_t(lp.cols)), lb(c.size()), ub(c.size()), rl(size_t(lp.rows)),
            ru(rl.size()), cs(c.size()), rs(rl.size());
        niryukti_options o;
        niryukti_default_options(&o);
        int64_t rows = 0, cols = 0;
        double osc = 0, off = 0;
        const int rc = niryukti_prepare_lp(&lp, &o, &rows, &cols, ptr.data(), index.data(), value.data(), c.data(),
                                           lb.data(), ub.data(), rl.data(), ru.data(), cs.data(), rs.data(),
                                           cmap.data(), rmap.data(), &osc, &off);
        expect(rc == 0 && rows <= lp.rows && cols <= lp.cols, "prepare");
        std::vector<double> x(size_t(lp.cols), 0.0), y(size_t(lp.rows), 0.0);
        niryukti_accuracy a{};
        expect(niryukti_verify_lp(&lp, x.data(), y.data(), 1, &a) == 0, "verify");
        niryukti_result r{};
        expect(niryukti_result_json(m, &r, "hip") != nullptr, "result document");
        niryukti_model_destroy(m);
    }
    expect(niryukti_model_load((tmp + "/missing.mps").c_str(), nullptr) == nullptr, "missing model refused");
    expect(niryukti_model_load(nullptr, "{\"variables\": 3}") == nullptr, "malformed JSON model refused");
}

void workloads() {
    const int configs[][3] = {{0, 20, 0}, {1, 400, 0}, {4, 200, 0}, {2, 1, 0}, {2, 1, 7}, {3, 1, 2}};
    for (const auto &k : configs) {
        niryukti_instance *p = niryukti_instance_create(k[0], 42, k[1], k[2]);
        expect(p != nullptr, "instance");
        if (!p) continue;
        niryukti_lp lp{};
        double planted = 0;
        const double *px = nullptr, *py = nullptr;
        niryukti_instance_view(p, &lp, &planted, &px, &py);
        niryukti_accuracy a{};
        expect(niryukti_verify_lp(&lp, px, py, 1, &a) == 0 && a.kkt <= 1e-9, "planted optimum certifies");
        expect(std::abs(a.objective - planted) <= 1e-9 * (1 + std::abs(planted)), "planted objective");
        niryukti_options o;
        niryukti_default_options(&o);
        niryukti_result r{};
        expect(niryukti_solve_lp(&lp, &o, &r) == 0 && std::string(r.status) == "UNSUPPORTED",
               "solve without a device is UNSUPPORTED");
        niryukti_instance_destroy(p);
    }
}

void session(const std::string &netlib) {
    niryukti_session *s = niryukti_open((netlib + "/afiro.mps").c_str());
    expect(s != nullptr, "session opens");
    if (!s) return;
    const char *out = niryukti_request(s, "{\"action\":\"solve\",\"tol\":1e-6}");
    expect(out && std::strstr(out, "UNSUPPORTED"), "session solve answers");
    expect(niryukti_request(s, "{\"action\":\"solve\",\"bogus\":1}") == nullptr, "unknown key refused");
    expect(niryukti_request(s, "not json") == nullptr, "bad JSON refused");
    niryukti_close(s);
}

void checkpoints(const std::string &netlib, const std::string &tmp) {
    niryukti_model *h = niryukti_model_load((netlib + "/afiro.mps").c_str(), nullptr);
    if (!h) return;
    niryukti_lp lp{};
    const char *name = nullptr, *fp = nullptr;
    niryukti_model_view(h, &lp, &name, &fp);
    nyk::Model m;
    m.A.rows = lp.rows;
    m.A.cols = lp.cols;
    m.A.ptr.assign(lp.ptr, lp.ptr + lp.rows + 1);
    m.A.index.assign(lp.index, lp.index + lp.ptr[lp.rows]);
    m.A.value.assign(lp.value, lp.value + lp.ptr[lp.rows]);
    m.c.assign(lp.c, lp.c + lp.cols);
    m.lb.assign(lp.lb, lp.lb + lp.cols);
    m.ub.assign(lp.ub, lp.ub + lp.cols);
    m.rl.assign(lp.rl, lp.rl + lp.rows);
    m.ru.assign(lp.ru, lp.ru + lp.rows);
    nyk::Options o;
    const nyk::Json config = nyk::checkpoint_configuration(o, false);
    nyk::LoopState s;
    s.backend = {{1.0, 2.0, nyk::kInf}, {-nyk::kInf, 0.5}};
    s.best_x.assign(size_t(lp.cols), 0.25);
    s.best_y.assign(size_t(lp.rows), -0.5);
    s.pid = {0.1, 0.2, 1.0};
    s.step = 0.9;
    s.weight = 2.0;
    s.restart_kkt = 0.01;
    s.iterations = 300;
    const std::string path = tmp + "/run.ckpt";
    nyk::write_checkpoint(path, m, config, s);
    const nyk::LoopState back = nyk::read_checkpoint(path, m, config);
    expect(back.backend == s.backend && back.best_x == s.best_x && back.iterations == 300 && back.step == 0.9,
           "checkpoint round trip");
    bool refused = false;
    try {
        nyk::read_checkpoint(path, m, nyk::checkpoint_configuration(o, true));
    } catch (const std::exception &) {
        refused = true;
    }
    expect(refused, "checkpoint of another configuration refused");
    niryukti_model_destroy(h);
}

// Many threads submit to one server at once; every request completes (with
// an error or UNSUPPORTED here) and every waiter returns.
void serving() {
    niryukti_server_limits l;
    niryukti_server_default_limits(&l);
    l.hbm_budget = int64_t(1) << 30;
    l.large_nonzeros = 3000;
    l.grid_workers = 3;
    niryukti_server *s = nullptr;
    expect(niryukti_server_create(nullptr, &l, &s) == 0, "server starts without a device");
    if (!s) return;
    std::vector<niryukti_instance *> insts;
    for (int i = 0; i < 8; ++i) insts.push_back(niryukti_instance_create(2, 7, 1, i));
    std::atomic<int> completed{0}, refused{0};
    std::vector<std::thread> clients;
    for (int t = 0; t < 4; ++t)
        clients.emplace_back([&, t] {
            for (int k = 0; k < 12; ++k) {
                niryukti_lp lp{};
                double planted = 0;
                const double *px = nullptr, *py = nullptr;
                niryukti_instance_view(insts[size_t((t * 12 + k) % 8)], &lp, &planted, &px, &py);
                std::vector<double> bad(size_t(lp.cols), 0.0);
                bad[0] = NAN;
                int64_t id = 0;
                if (k % 5 == 4) {
                    if (niryukti_server_submit(s, &lp, bad.data(), nullptr, -1, nullptr, &id) != 0) ++refused;
                    continue;
                }
                if (niryukti_server_submit(s, &lp, px, py, k % 3 == 0 ? 1 : -1, nullptr, &id) != 0) continue;
                std::vector<double> x(size_t(lp.cols)), y(size_t(lp.rows));
                niryukti_result r{};
                r.x = x.data();
                r.y = y.data();
                double latency = 0;
                int route = -1;
                const int rc = niryukti_server_wait(s, id, -1.0, &r, &latency, &route);
                if (rc == 1 || rc == -1) ++completed;
                niryukti_server_release(s, id);
            }
        });
    for (auto &c : clients) c.join();
    niryukti_server_stats st{};
    expect(niryukti_server_stats_get(s, &st) == 0 && st.finished == st.submitted, "every request finished");
    expect(completed == 4 * 12 - refused && refused > 0, "every waiter returned; bad warm starts refused");
    niryukti_server_destroy(s);
    for (auto *p : insts) niryukti_instance_destroy(p);
}

// The batched engine's certification beside the next round: a task checks a
// set of candidate points with the verifier on its own thread while the
// caller keeps working on other data, then the caller applies the outcome.
void certification_beside() {
    std::vector<niryukti_instance *> insts;
    for (int i = 0; i < 6; ++i) insts.push_back(niryukti_instance_create(2, 11, 1, i));
    struct Record {
        nyk::Model model;
        std::vector<double> x, y;
        double kkt = -1;
    };
    std::vector<Record> records(insts.size());
    for (size_t k = 0; k < insts.size(); ++k) {
        niryukti_lp lp{};
        double planted = 0;
        const double *px = nullptr, *py = nullptr;
        niryukti_instance_view(insts[k], &lp, &planted, &px, &py);
        nyk::Model &m = records[k].model;
        m.A.rows = lp.rows;
        m.A.cols = lp.cols;
        m.A.ptr.assign(lp.ptr, lp.ptr + lp.rows + 1);
        m.A.index.assign(lp.index, lp.index + lp.ptr[lp.rows]);
        m.A.value.assign(lp.value, lp.value + lp.ptr[lp.rows]);
        m.c.assign(lp.c, lp.c + lp.cols);
        m.lb.assign(lp.lb, lp.lb + lp.cols);
        m.ub.assign(lp.ub, lp.ub + lp.cols);
        m.rl.assign(lp.rl, lp.rl + lp.rows);
        m.ru.assign(lp.ru, lp.ru + lp.rows);
        records[k].x.assign(px, px + lp.cols);
        records[k].y.assign(py, py + lp.rows);
    }
    for (int round = 0; round < 3; ++round) {
        nyk::Beside task;
        task.start([&] {
            nyk::parallel_tasks(int64_t(records.size()), [&](int64_t k) {
                nyk::Verifier v(records[size_t(k)].model);
                records[size_t(k)].kkt = v.evaluate(records[size_t(k)].x, records[size_t(k)].y).kkt;
            });
        });
        // the caller's own work meanwhile, on data the task does not touch
        std::vector<double> other(100000, 1.0);
        double sum = 0;
        for (double v : other) sum += v;
        expect(sum == 100000.0, "caller work");
        task.wait();
        for (const auto &r : records) expect(r.kkt >= 0 && r.kkt <= 1e-9, "certified beside");
    }
    nyk::Beside failing;
    failing.start([] { throw std::runtime_error("certification failed"); });
    bool rethrown = false;
    try {
        failing.wait();
    } catch (const std::runtime_error &) {
        rethrown = true;
    }
    expect(rethrown, "a failing task's error reaches the caller");
    for (auto *p : insts) niryukti_instance_destroy(p);
}
// The batched engine's submission inbox and job reaper under concurrent
// producers: every item gets a distinct sequence number equal to its position
// in the consumer's list, while the consumer drains beside the producers, and
// every item given to the reaper is destroyed exactly once.
void inbox_and_reaper() {
    struct Item {
        int64_t ticket = -1, producer = 0, seq = 0;
        std::vector<double> payload;
    };
    nyk::Inbox<std::unique_ptr<Item>> inbox;
    constexpr int kProducers = 4, kEach = 2000;
    std::vector<std::thread> producers;
    std::vector<std::vector<int64_t>> returned(kProducers);
    for (int w = 0; w < kProducers; ++w)
        producers.emplace_back([&, w] {
            for (int k = 0; k < kEach; ++k) {
                auto item = std::make_unique<Item>();
                item->producer = w;
                item->seq = k;
                item->payload.assign(64, double(k));
                returned[size_t(w)].push_back(
                    inbox.push(std::move(item), [](std::unique_ptr<Item> &q, int64_t n) { q->ticket = n; }));
            }
        });
    std::vector<std::unique_ptr<Item>> filed;
    while (filed.size() < size_t(kProducers * kEach)) inbox.drain(filed);  // beside the producers
    for (auto &t : producers) t.join();
    inbox.drain(filed);
    expect(filed.size() == size_t(kProducers * kEach) && inbox.pushed() == kProducers * kEach, "inbox keeps every item");
    bool at_own_index = true;
    std::vector<int64_t> last(kProducers, -1);
    for (size_t k = 0; k < filed.size(); ++k) {
        at_own_index = at_own_index && filed[k]->ticket == int64_t(k);
        // one producer's items keep their order
        at_own_index = at_own_index && filed[k]->seq > last[size_t(filed[k]->producer)];
        last[size_t(filed[k]->producer)] = filed[k]->seq;
    }
    expect(at_own_index, "sequence numbers are positions");
    for (int w = 0; w < kProducers; ++w)
        for (int k = 0; k < kEach; ++k)
            expect(filed[size_t(returned[size_t(w)][size_t(k)])]->seq == k, "push returns the item's number");

    std::atomic<int64_t> destroyed{0};
    struct Counted {
        explicit Counted(std::atomic<int64_t> *c) : count(c), payload(512, 1.0) {}
        Counted(const Counted &) = delete;
        ~Counted() { count->fetch_add(1); }
        std::atomic<int64_t> *count;
        std::vector<double> payload;
    };
    {
        nyk::Reaper<std::unique_ptr<Counted>> reaper;
        std::vector<std::thread> givers;
        for (int w = 0; w < 2; ++w)
            givers.emplace_back([&] {
                for (int k = 0; k < 5000; ++k)
                    reaper.give(std::make_unique<Counted>(&destroyed));
            });
        for (auto &t : givers) t.join();
    }  // the destructor destroys what is left and joins
    expect(destroyed.load() == 10000, "the reaper destroys every item once");
}
}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: harness NETLIB_DIR TMP_DIR\n");
        return 2;
    }
    const std::string netlib = argv[1], tmp = argv[2];
    try {
        models(netlib, tmp);
        workloads();
        session(netlib);
        checkpoints(netlib, tmp);
        serving();
        certification_beside();
        inbox_and_reaper();
    } catch (const std::exception &e) {
        std::fprintf(stderr, "FAIL: exception %s\n", e.what());
        return 1;
    }
    std::printf("harness: %d failure(s)\n", failures);
    return failures ? 1 : 0;
}
