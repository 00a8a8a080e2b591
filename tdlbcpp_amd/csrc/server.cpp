// Serving loop for a stream of tank requests (no counterpart in the reference).
//
// Requests (grid edge, tip speed, Reynolds number, step count) queue in arrival
// order.  Each tick the server
//   1. admits queued requests in order while their HBM footprint fits in the
//      budget next to the resident tanks (a request that does not fit waits,
//      and so do the ones behind it);
//   2. hands admitted requests to a preparation thread, which allocates,
//      initialises and binds the tank while the GPU keeps running;
//   3. adds prepared tanks to the batch (continuous batching: tanks join and
//      leave between ticks, never waiting for the others to finish);
//   4. runs the batch for tick_steps steps;
//   5. retires tanks whose steps are done: records the result (FP64 mass and
//      power number from the device sums), frees their memory.
// A tank in the batch advances bit for bit as it would alone (tdlb_batch_*).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/tdlb.h"
#include "handle.h"
#include "rushton.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    int64_t ticket = 0;
    tdlb_tank_request req{};
    double submit_s = 0, admit_s = 0, start_s = 0;
    size_t footprint = 0;
    tdlb_cu_t cu = nullptr;
    tdlb_rushton_params rp{};
    double mass0 = 0;
    int error = TDLB_OK;
    std::string error_msg;
};

}  // namespace

struct tdlb_server {
    int prec = 0, layout = 0, streaming = 0, device = 0;
    double budget = 0;
    int64_t tick_steps = 50;
    int keep_handles = 0;
    int64_t max_resident = 0;  // 0: no limit besides memory
    Clock::time_point t0;
    tdlb_batch_t batch = nullptr;

    int64_t next_ticket = 1;
    std::deque<std::unique_ptr<Job>> queue;      // submitted, not admitted
    std::vector<std::unique_ptr<Job>> running;   // in the batch
    std::vector<tdlb_tank_result> done;          // finished, not yet polled
    std::vector<std::pair<int64_t, tdlb_cu_t>> kept;
    size_t reserved = 0;                          // footprint of admitted tanks

    // preparation thread
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::unique_ptr<Job>> to_prepare, prepared;
    int preparing = 0;
    bool stop = false;
    std::thread worker;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

// bytes a tank holds on the device: lattices, F, Nu, O and the masks, plus 10 %
// for the IB lists, level tables and caches
size_t footprint(const tdlb_server* s, int64_t edge) {
    const size_t n = (size_t)(edge + 2) * (edge + 2) * (edge + 2);
    const size_t pop = s->prec == TDLB_PREC_F64 ? 8 : s->prec == TDLB_PREC_F32 ? 4 : 2;
    const size_t t = s->prec == TDLB_PREC_F64 ? 8 : 4;
    const size_t nq = s->streaming == TDLB_STREAM_SIMPLE ? 2 : 1;
    // population cells as the handle lays them out
    const int32_t zg = (int32_t)edge + 2;
    Grid g{};
    g.xg = g.yg = g.zg = zg;
    g.xg1 = g.yg1 = g.zg1 = zg - 2;
    g.yz = (uint32_t)(zg * zg);
    g.size = n;
    set_q_geometry(g, s->layout == TDLB_LAYOUT_IJKL ? LayoutIJKL : LayoutLIJK, pop);
    const size_t qcells = (size_t)g.qsize;
    const size_t masks = 5 * (size_t)zg * zg * ((zg + 63) / 64) * 8;
    return (size_t)((double)(nq * 18 * pop * qcells + n * (3 * t + (pop == 2 ? 2 : t) + 1) + masks) * 1.10);
}

// The tank of a request: FlowParams as main.cpp leaves them (nu from
// calcNuAndRe_m with the integer blade radius), the generator's default tank at
// the request's tip speed, initialised at rest (rho = 1) and bound at step 1.
void prepare(tdlb_server* s, Job& j) {
    const int64_t T = j.req.edge;
    tdlb_flow f{};
    f.initial_rho = 1.0;
    f.re_m_nondimensional = j.req.reynolds;
    f.uav = j.req.tip_speed;
    f.cs0 = 0.1;
    f.g3 = 0.8;
    f.alpha = 0.97;
    f.beta = 1.9;
    f.use_les = 1;
    f.collision = 0;
    int rc = tdlb_flow_calc_nu_re_m(&f, (int32_t)(T / 6), s->prec);
    tdlb_cu_params p{};
    p.x = p.y = p.z = T;
    p.ghost = 1;
    p.resolution = 1;
    p.device_id = s->device;
    p.layout = s->layout;
    p.streaming = s->streaming;
    p.precision = s->prec;
    if (rc == TDLB_OK) rc = tdlb_cu_create(&p, &f, &j.cu);
    // the tile timings of a new grid shape run here, on the tank's own arrays
    // before they are initialised: no scratch lattice, nothing on the tick thread
    if (rc == TDLB_OK) rc = tdlb_batch_tune(s->batch, j.cu);
    if (rc == TDLB_OK) rc = tdlb_cu_initialise(j.cu, 1.0);
    if (rc == TDLB_OK) rc = tdlb_rushton_defaults(T, T, j.req.tip_speed, j.req.startup_steps, &j.rp);
    if (rc == TDLB_OK) rc = tdlb_cu_bind_rushton(j.cu, &j.rp, 1, 0.0, f.alpha, f.beta);
    double red[8];
    if (rc == TDLB_OK) rc = tdlb_cu_reductions(j.cu, red);
    if (rc == TDLB_OK) j.mass0 = red[0];
    if (rc != TDLB_OK) {
        j.error = rc;
        j.error_msg = tdlb_last_error();
        if (j.cu) tdlb_cu_destroy(j.cu);
        j.cu = nullptr;
    }
}

void worker_main(tdlb_server* s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv.wait(lk, [&] { return s->stop || !s->to_prepare.empty(); });
        if (s->stop) return;
        std::unique_ptr<Job> j = std::move(s->to_prepare.front());
        s->to_prepare.pop_front();
        lk.unlock();
        prepare(s, *j);
        lk.lock();
        s->prepared.push_back(std::move(j));
        s->preparing--;
        s->cv.notify_all();
    }
}

tdlb_tank_result result_of(const tdlb_server* s, const Job& j, double finish, int error, const double* red,
                           double dtheta) {
    tdlb_tank_result r{};
    r.ticket = j.ticket;
    r.edge = j.req.edge;
    r.steps = j.req.steps;
    r.error = error;
    r.submit_s = j.submit_s;
    r.admit_s = j.admit_s;
    r.start_s = j.start_s;
    r.finish_s = finish;
    if (red) {
        r.mass = red[0];
        r.mass_drift = j.mass0 != 0 ? red[0] / j.mass0 - 1.0 : 0.0;
        r.torque = red[7];
        const double D = 2.0 * j.rp.blade_outer_radius;
        const double N = dtheta / (2.0 * M_PI);
        r.power_number = (N != 0 && D > 0) ? red[7] * dtheta / (1.0 * N * N * N * std::pow(D, 5)) : 0.0;
    }
    (void)s;
    return r;
}

}  // namespace

extern "C" {

int tdlb_server_create(int32_t precision, int32_t layout, int32_t streaming, int32_t device_id,
                       double hbm_budget_bytes, int64_t tick_steps, int64_t max_resident, int32_t keep_handles,
                       tdlb_server_t* out) {
    if (!out || tick_steps < 1) return fail(TDLB_ERR_INVALID, "bad arguments");
    tdlb_server* s = new tdlb_server();
    s->prec = precision;
    s->layout = layout;
    s->streaming = streaming;
    s->device = device_id < 0 ? 0 : device_id;
    s->tick_steps = tick_steps;
    s->keep_handles = keep_handles;
    s->max_resident = max_resident < 0 ? 0 : max_resident;
    s->t0 = Clock::now();
    int rc = tdlb_batch_create(precision, layout, streaming, s->device, &s->batch);
    if (rc != TDLB_OK) {
        delete s;
        return rc;
    }
    size_t freeb = 0, total = 0;
    if (hipSetDevice(s->device) != hipSuccess || hipMemGetInfo(&freeb, &total) != hipSuccess) {
        tdlb_batch_destroy(s->batch);
        delete s;
        return fail(TDLB_ERR_HIP, "device memory query");
    }
    // the budget is capped at 90 % of the memory free now
    const double cap = 0.9 * (double)freeb;
    s->budget = hbm_budget_bytes > 0 ? std::min(hbm_budget_bytes, cap) : cap;
    s->worker = std::thread(worker_main, s);
    *out = s;
    return TDLB_OK;
}

int tdlb_server_destroy(tdlb_server_t s) {
    if (!s) return TDLB_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    auto drop = [](Job& j) {
        if (j.cu) tdlb_cu_destroy(j.cu);
        j.cu = nullptr;
    };
    for (auto& j : s->running) drop(*j);
    for (auto& j : s->prepared) drop(*j);
    for (auto& j : s->to_prepare) drop(*j);
    for (auto& k : s->kept) tdlb_cu_destroy(k.second);
    tdlb_batch_destroy(s->batch);
    delete s;
    return TDLB_OK;
}

int tdlb_server_submit(tdlb_server_t s, const tdlb_tank_request* req, int64_t* ticket) {
    if (!s || !req) return fail(TDLB_ERR_INVALID, "null argument");
    if (req->edge < 12 || req->steps < 1 || req->tip_speed < 0 || req->reynolds <= 0)
        return fail(TDLB_ERR_INVALID, "request: edge >= 12, steps >= 1, tip_speed >= 0, reynolds > 0");
    auto j = std::make_unique<Job>();
    j->ticket = s->next_ticket++;
    j->req = *req;
    j->submit_s = s->now();
    j->footprint = footprint(s, req->edge);
    if ((double)j->footprint > s->budget) return fail(TDLB_ERR_NOMEM, "request larger than the HBM budget");
    if (ticket) *ticket = j->ticket;
    s->queue.push_back(std::move(j));
    return TDLB_OK;
}

int tdlb_server_tick(tdlb_server_t s, int64_t* resident, int64_t* queued, int64_t* steps_run) {
    if (!s) return fail(TDLB_ERR_INVALID, "null server");
    if (steps_run) *steps_run = 0;
    // 1-2. admission by footprint, in arrival order
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto admitted = [&]() {
            return (int64_t)(s->running.size() + s->prepared.size()) + s->preparing;
        };
        while (!s->queue.empty() && (double)(s->reserved + s->queue.front()->footprint) <= s->budget &&
               (s->max_resident == 0 || admitted() < s->max_resident)) {
            std::unique_ptr<Job> j = std::move(s->queue.front());
            s->queue.pop_front();
            s->reserved += j->footprint;
            j->admit_s = s->now();
            s->to_prepare.push_back(std::move(j));
            s->preparing++;
        }
    }
    s->cv.notify_all();
    // 3. join prepared tanks; with nothing to run, wait for the first one
    {
        std::unique_lock<std::mutex> lk(s->mu);
        if (s->running.empty() && s->prepared.empty() && s->preparing > 0)
            s->cv.wait(lk, [&] { return !s->prepared.empty() || s->preparing == 0; });
        while (!s->prepared.empty()) {
            std::unique_ptr<Job> j = std::move(s->prepared.front());
            s->prepared.pop_front();
            if (j->error != TDLB_OK) {
                s->reserved -= j->footprint;
                s->done.push_back(result_of(s, *j, s->now(), j->error, nullptr, 0.0));
                continue;
            }
            lk.unlock();
            const int rc = tdlb_batch_add(s->batch, j->cu, 1, 0.0, j->req.steps);
            lk.lock();
            if (rc != TDLB_OK) {
                tdlb_cu_destroy(j->cu);
                j->cu = nullptr;
                s->reserved -= j->footprint;
                s->done.push_back(result_of(s, *j, s->now(), rc, nullptr, 0.0));
                continue;
            }
            j->start_s = s->now();
            s->running.push_back(std::move(j));
        }
    }
    // 4. one tick of the batch
    if (!s->running.empty()) {
        int64_t n = 0;
        CHK(tdlb_batch_run(s->batch, s->tick_steps, &n));
        if (steps_run) *steps_run = n;
    }
    // 5. retire finished tanks
    for (size_t a = 0; a < s->running.size();) {
        Job& j = *s->running[a];
        uint64_t step = 0;
        double angle = 0;
        int64_t rem = 0;
        CHK(tdlb_batch_member(s->batch, j.cu, &step, &angle, &rem));
        if (rem > 0) {
            a++;
            continue;
        }
        double red[8] = {0};
        const int rc = tdlb_cu_reductions(j.cu, red);
        double dth = 0;
        (void)tdlb_rushton_step_increment(&j.rp, step, &dth);
        if (j.cu->tsize == 4) dth = (double)(float)dth;
        CHK(tdlb_batch_remove(s->batch, j.cu));
        s->done.push_back(result_of(s, j, s->now(), rc, rc == TDLB_OK ? red : nullptr, dth));
        if (s->keep_handles) s->kept.push_back({j.ticket, j.cu});
        else tdlb_cu_destroy(j.cu);
        j.cu = nullptr;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->reserved -= j.footprint;
        }
        s->running.erase(s->running.begin() + (long)a);
    }
    if (resident) *resident = (int64_t)s->running.size();
    if (queued) {
        std::lock_guard<std::mutex> lk(s->mu);
        *queued = (int64_t)(s->queue.size() + s->prepared.size()) + s->preparing;
    }
    return TDLB_OK;
}

int tdlb_server_poll(tdlb_server_t s, tdlb_tank_result* out, int64_t cap, int64_t* n) {
    if (!s || !n) return fail(TDLB_ERR_INVALID, "null argument");
    const int64_t m = std::min<int64_t>(cap, (int64_t)s->done.size());
    for (int64_t a = 0; a < m; a++) out[a] = s->done[a];
    s->done.erase(s->done.begin(), s->done.begin() + m);
    *n = m;
    return TDLB_OK;
}

int tdlb_server_take(tdlb_server_t s, int64_t ticket, tdlb_cu_t* cu) {
    if (!s || !cu) return fail(TDLB_ERR_INVALID, "null argument");
    for (size_t a = 0; a < s->kept.size(); a++)
        if (s->kept[a].first == ticket) {
            *cu = s->kept[a].second;
            s->kept.erase(s->kept.begin() + (long)a);
            return TDLB_OK;
        }
    return fail(TDLB_ERR_INVALID, "no kept handle for this ticket");
}

int tdlb_server_stats(tdlb_server_t s, double* now_s, double* reserved_bytes, double* budget_bytes) {
    if (!s) return fail(TDLB_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (now_s) *now_s = s->now();
    if (reserved_bytes) *reserved_bytes = (double)s->reserved;
    if (budget_bytes) *budget_bytes = s->budget;
    return TDLB_OK;
}

}  // extern "C"
