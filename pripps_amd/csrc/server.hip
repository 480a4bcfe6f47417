// Serving loop for Debye solves: continuous batching with admission by free
// device memory.
//
// Requests are copied at submit time and queued.  Three worker lanes, each with
// its own engine (stream + device arena) on the same GPU, repeatedly take the
// oldest queued requests whose estimated device footprint fits what is left
// of the memory budget, and build them as one batched persistent launch.  A
// lane that is building leaves the queue to the other lanes, so requests that
// arrive while a batch is on the device start in the next batch and run
// alongside it (the second kernel's waves take CUs as the first one's drain).
// Finished models keep their profiles on the device and are handed to the
// caller by pripps_server_wait together with the request's latency (submit to
// completion).  They move from their lane to the engine the server was
// created on, so that the caller's intensity() and fit() calls on them do not
// wait for the lane's next batch.
//
// The reference has no serving layer; this loop wraps the build that replaces
// debye::build_model (src/debye.rs:78-136).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "internal.h"
#include "pripps_debye.h"

namespace pripps {
int engine_build(pripps_engine *e, int32_t n, const pripps_debye_input *inputs, int32_t precision,
                 const pripps_tuning *tuning, pripps_debye_model **out);
void set_last_error(const std::string &msg);
const std::string &get_last_error();
void model_rehome(pripps_debye_model *m, pripps_engine *to);
bool engine_estimate(pripps_engine *e, const pripps_debye_input *in, int32_t precision,
                     uint64_t *arena_bytes, uint64_t *model_block);
uint64_t engine_arena_bytes(pripps_engine *e);
int engine_trim_arena(pripps_engine *e);
void engine_retain(pripps_engine *e);
void engine_unref(pripps_engine *e);
}  // namespace pripps

using namespace pripps;
using Clock = std::chrono::steady_clock;

namespace {

struct Request {
    int64_t id = 0;
    Clock::time_point t_submit;
    // owned copies of the caller's arrays
    std::vector<double> x, y, z, q, ffa, ffe, ffw, w;
    std::vector<int32_t> kind;
    pripps_debye_input in{};
    uint64_t arena = 0;  // device bytes of its plan's arena when built alone
    uint64_t model_bytes = 0;  // device bytes of its model block
    // completion
    bool done = false;
    bool waiting = false;  // a pripps_server_wait call holds this request
    int status = PRIPPS_OK;
    std::string error;
    pripps_debye_model *model = nullptr;
    double latency_ms = 0;
};

// Worker lanes, each an engine (stream + arena) with a thread that prepares,
// uploads and launches its batch and waits for it.  Three keep a pair kernel
// on the device while the other lanes are between batches (measured at 1,000
// offered requests/s: 1.19e13 terms/s with two lanes, 1.22e13 with three, no
// gain from a fourth).
constexpr int kLanes = 3;

// Fallback footprint of an input that does not prepare (its build fails at
// validation and needs no arena): the model block.
uint64_t fallback_bytes(const pripps_debye_input &in) {
    return (uint64_t)(11.0 * (double)std::max<int32_t>(in.nq, 1) * 8.0) + 4096;
}

}  // namespace

struct pripps_server {
    int device = 0;
    pripps_engine *base = nullptr;  // the caller's engine: finished models answer queries on it
    pripps_server_config cfg{};
    std::vector<pripps_engine *> lanes;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;
    std::unordered_map<int64_t, Request *> all;
    int64_t next_id = 1;
    // Admission.  `budget` is fixed when the server is created: the
    // configured figure, or 90 % of the device memory free at that moment.
    // Charged against it: every lane's arena (grow-only, it stays allocated
    // between batches) or, while the lane builds, the larger of that arena
    // and the arena its batch needs, plus the batch's model blocks until they
    // are handed over.  Models the caller holds after pripps_server_wait are
    // the caller's memory, outside the budget.  A request whose own arena and
    // model exceed the budget can never run and fails at once with
    // PRIPPS_ERR_OUT_OF_MEMORY; any other waits until it fits, and idle lanes
    // give their arenas back if that is what keeps it out.
    uint64_t budget = 0;
    std::vector<uint64_t> lane_need;    // arena bytes the lane's current batch needs (0: idle)
    std::vector<uint64_t> lane_models;  // model bytes of the lane's current batch
    bool stopping = false;

    // device bytes charged to lane k (caller holds mu)
    uint64_t charged(int k) const {
        return std::max(engine_arena_bytes(lanes[k]), lane_need[k]) + lane_models[k];
    }

    void worker(int me) {
        pripps_engine *lane = lanes[me];
        for (;;) {
            std::vector<Request *> batch;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stopping || !queue.empty(); });
                if (stopping && queue.empty()) return;
                const int cap = cfg.max_in_flight > 0 ? cfg.max_in_flight : 256;
                uint64_t others = 0, idle_held = 0;
                for (int k = 0; k < (int)lanes.size(); ++k) {
                    if (k == me) continue;
                    others += charged(k);
                    if (lane_need[k] == 0) idle_held += engine_arena_bytes(lanes[k]);
                }
                const uint64_t own = engine_arena_bytes(lane);
                idle_held += own;
                uint64_t need = 0, models = 0;
                bool head_blocked = false;
                while (!queue.empty() && (int)batch.size() < cap) {
                    Request *r = queue.front();
                    if (r->arena + r->model_bytes > budget) {
                        // can never fit: fail it rather than stall the queue
                        queue.pop_front();
                        r->done = true;
                        r->status = PRIPPS_ERR_OUT_OF_MEMORY;
                        r->error = "request needs about " + std::to_string(r->arena + r->model_bytes) +
                                   " bytes of device memory; the serving budget is " +
                                   std::to_string(budget);
                        r->latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - r->t_submit).count();
                        cv_done.notify_all();
                        continue;
                    }
                    const uint64_t mine = std::max(own, need + r->arena) + models + r->model_bytes;
                    if (others + mine > budget) {  // wait for memory
                        head_blocked = batch.empty();
                        break;
                    }
                    queue.pop_front();
                    batch.push_back(r);
                    need += r->arena;
                    models += r->model_bytes;
                }
                if (batch.empty()) {
                    if (head_blocked && idle_held > 0) {
                        // idle lanes' arenas (this one's included) keep the
                        // head out: give them back
                        for (int k = 0; k < (int)lanes.size(); ++k)
                            if (k == me || lane_need[k] == 0) (void)engine_trim_arena(lanes[k]);
                        continue;
                    }
                    // the head does not fit while other batches hold memory
                    cv_work.wait_for(lk, std::chrono::milliseconds(1));
                    continue;
                }
                lane_need[me] = std::max<uint64_t>(need, 1);
                lane_models[me] = models;
            }
            const size_t nb = batch.size();
            std::vector<pripps_debye_input> ins(nb);
            for (size_t k = 0; k < nb; ++k) ins[k] = batch[k]->in;
            std::vector<pripps_debye_model *> out(nb, nullptr);
            std::vector<int> rcs(nb, PRIPPS_OK);
            std::vector<std::string> errs(nb);
            const int rc = engine_build(lane, (int32_t)nb, ins.data(), cfg.precision, nullptr,
                                        out.data());
            if (rc != PRIPPS_OK) {
                // A batch fails as a whole (one request with non-finite input
                // fails its validation).  Build each request alone so that
                // only the requests that fail by themselves report an error.
                const std::string err = get_last_error();
                for (size_t k = 0; k < nb; ++k) {
                    out[k] = nullptr;
                    rcs[k] = nb == 1 ? rc : engine_build(lane, 1, &ins[k], cfg.precision, nullptr, &out[k]);
                    if (rcs[k] != PRIPPS_OK) errs[k] = nb == 1 ? err : get_last_error();
                }
            }
            for (size_t k = 0; k < nb; ++k)
                if (out[k]) model_rehome(out[k], base);
            const auto now = Clock::now();
            {
                std::lock_guard<std::mutex> lk(mu);
                lane_need[me] = 0;
                lane_models[me] = 0;
                for (size_t k = 0; k < nb; ++k) {
                    Request *r = batch[k];
                    r->done = true;
                    r->status = rcs[k];
                    r->error = errs[k];
                    r->model = out[k];
                    r->latency_ms = std::chrono::duration<double, std::milli>(now - r->t_submit).count();
                    // inputs are no longer needed
                    r->x.clear(); r->x.shrink_to_fit();
                    r->y.clear(); r->y.shrink_to_fit();
                    r->z.clear(); r->z.shrink_to_fit();
                    r->kind.clear(); r->kind.shrink_to_fit();
                    r->w.clear(); r->w.shrink_to_fit();
                }
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" int pripps_server_create(pripps_engine *engine, const pripps_server_config *config,
                                    pripps_server **out) {
    if (!engine || !out) {
        set_last_error("pripps_server_create: NULL argument");
        return PRIPPS_ERR_USAGE;
    }
    *out = nullptr;
    uint64_t f = 0, t = 0;
    int rc = pripps_engine_memory(engine, &f, &t);
    if (rc) return rc;
    auto *s = new pripps_server();
    if (config) s->cfg = *config;
    s->budget = s->cfg.memory_budget ? s->cfg.memory_budget : (uint64_t)(0.9 * (double)f);
    if (s->cfg.precision != PRIPPS_FP32_COMPENSATED && s->cfg.precision != PRIPPS_FP64) {
        delete s;
        set_last_error("pripps_server_create: precision must be FP32_COMPENSATED or FP64");
        return PRIPPS_ERR_USAGE;
    }
    // pripps_engine_memory made the engine's device current
    int dev = 0;
    (void)hipGetDevice(&dev);
    s->device = dev;
    for (int k = 0; k < kLanes; ++k) {
        pripps_engine *lane = nullptr;
        rc = pripps_engine_create(dev, &lane);
        if (rc) {
            for (auto *l : s->lanes) pripps_engine_destroy(l);
            delete s;
            return rc;
        }
        s->lanes.push_back(lane);
    }
    engine_retain(engine);  // released by pripps_server_destroy
    s->base = engine;
    s->lane_need.assign(kLanes, 0);
    s->lane_models.assign(kLanes, 0);
    for (int k = 0; k < kLanes; ++k) s->workers.emplace_back([s, k] { s->worker(k); });
    *out = s;
    return PRIPPS_OK;
}

extern "C" int pripps_server_submit(pripps_server *s, const pripps_debye_input *in,
                                    int64_t *request_id) {
    if (!s || !in || !request_id) {
        set_last_error("pripps_server_submit: NULL argument");
        return PRIPPS_ERR_USAGE;
    }
    if (in->n_atoms < 0 || in->nq < 0 || in->n_kinds < 1 ||
        (in->n_atoms > 0 && (!in->x || !in->y || !in->z || !in->kind)) ||
        (in->nq > 0 && (!in->q || !in->atom_ff)) ||
        ((in->water_ff != nullptr) != (in->hydration_weight != nullptr))) {
        set_last_error("pripps_server_submit: invalid input descriptor");
        return PRIPPS_ERR_USAGE;
    }
    auto *r = new Request();
    const size_t n = (size_t)in->n_atoms, nq = (size_t)in->nq, kq = (size_t)in->n_kinds * nq;
    r->x.assign(in->x, in->x + n);
    r->y.assign(in->y, in->y + n);
    r->z.assign(in->z, in->z + n);
    r->kind.assign(in->kind, in->kind + n);
    r->q.assign(in->q, in->q + nq);
    r->ffa.assign(in->atom_ff, in->atom_ff + kq);
    if (in->excluded_ff) r->ffe.assign(in->excluded_ff, in->excluded_ff + kq);
    if (in->water_ff) {
        r->ffw.assign(in->water_ff, in->water_ff + nq);
        r->w.assign(in->hydration_weight, in->hydration_weight + n);
    }
    r->in = *in;
    r->in.x = r->x.data();
    r->in.y = r->y.data();
    r->in.z = r->z.data();
    r->in.kind = r->kind.data();
    r->in.q = r->q.data();
    r->in.atom_ff = r->ffa.data();
    r->in.excluded_ff = in->excluded_ff ? r->ffe.data() : nullptr;
    r->in.water_ff = in->water_ff ? r->ffw.data() : nullptr;
    r->in.hydration_weight = in->water_ff ? r->w.data() : nullptr;
    // footprint from the plan the build will use (outside the server lock)
    if (!engine_estimate(s->lanes[0], &r->in, s->cfg.precision, &r->arena, &r->model_bytes)) {
        r->arena = 0;
        r->model_bytes = fallback_bytes(*in);
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        r->t_submit = Clock::now();
        s->all[r->id] = r;
        s->queue.push_back(r);
        *request_id = r->id;
    }
    s->cv_work.notify_one();
    return PRIPPS_OK;
}

extern "C" int pripps_server_wait(pripps_server *s, int64_t id, int32_t timeout_ms,
                                  pripps_debye_model **model, double *latency_ms) {
    if (!s || !model) {
        set_last_error("pripps_server_wait: NULL argument");
        return PRIPPS_ERR_USAGE;
    }
    *model = nullptr;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) {
        set_last_error("pripps_server_wait: unknown request id " + std::to_string(id));
        return PRIPPS_ERR_USAGE;
    }
    Request *r = it->second;
    if (r->waiting) {
        // the other waiter frees the request when it returns
        set_last_error("pripps_server_wait: request " + std::to_string(id) +
                       " is already being waited for");
        return PRIPPS_ERR_USAGE;
    }
    r->waiting = true;
    auto ready = [&] { return r->done; };
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
        r->waiting = false;
        set_last_error("request " + std::to_string(id) + " still running");
        return PRIPPS_ERR_NOT_READY;
    }
    // by key: submits during the wait may have rehashed the map, which
    // invalidates the iterator found above
    s->all.erase(id);
    const int status = r->status;
    const std::string err = r->error;
    *model = r->model;
    if (latency_ms) *latency_ms = r->latency_ms;
    delete r;
    lk.unlock();
    if (status) set_last_error(err);
    return status;
}

extern "C" int pripps_server_destroy(pripps_server *s) {
    if (!s) return PRIPPS_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    for (auto &th : s->workers) th.join();
    for (auto &kv : s->all) {
        if (kv.second->model) pripps_debye_model_free(kv.second->model);
        delete kv.second;
    }
    for (auto *l : s->lanes) pripps_engine_destroy(l);
    engine_unref(s->base);
    delete s;
    return PRIPPS_OK;
}
