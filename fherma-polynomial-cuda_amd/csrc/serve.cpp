// Batched and served forms: many independent products, each with its own point
// (N, W, L, q), through a fixed set of worker lanes and one device memory arena.
// No counterpart in the reference, whose fherma_run multiplies one pair at a
// time (candidates/ordered-streaming/solve.cu:1226-1271).
//
// A request's life: a planner thread builds the host part of its plan (prime
// walk, M_k; tables come later on the device) as soon as it arrives, up to
// 2 * lanes requests ahead of the workers, so plan construction overlaps the
// products in flight.  A worker then takes the oldest request, reserves its
// plan's exact device footprint in the arena -- waiting there while other
// requests hold the memory, which is the admission rule -- builds the device
// tables, runs the host-vector product on the lane's streams, and gives the
// block back.  Requests on different lanes overlap on the device (transfers of
// one with kernels of another, small products side by side); the lane's
// streams, events, pinned staging and copy workers are made once and reused by
// every request it runs.
#include <algorithm>
#include <chrono>
#include <deque>
#include <exception>
#include <unordered_map>

#include "polymul.h"
#include "state.hpp"

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Job {
    pm_request request;
    Clock::time_point submitted, admitted, finished_at;
    pm_status status = PM_OK;
    std::string error;
    bool finished = false;
    // built ahead by a planner (guarded by the server's mutex until planned)
    std::unique_ptr<pm::HostPlan> host;
    std::exception_ptr plan_error;
    bool planned = false, taken = false;
    bool claimed = false;  // a pm_server_wait holds this ticket (guarded by the server's mutex)
};

constexpr unsigned kPlanners = 2;

}  // namespace

struct pm_server {
    int device_id = 0;
    std::unique_ptr<pm::DeviceArena> arena;
    std::vector<std::unique_ptr<pm::Lane>> lanes;
    std::vector<std::thread> workers, planners;
    std::mutex mutex;
    std::condition_variable work, done, plan_work, planned;
    std::deque<Job*> queue, plan_queue;
    size_t planned_ahead = 0, lookahead = 2;  // plans built for requests no worker has taken
    std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
    uint64_t next_ticket = 1;
    bool stopping = false;

    void planner() {
        const unsigned hw = std::thread::hardware_concurrency();
        pm::set_plan_threads(std::max(2u, (hw ? hw : 4u) / kPlanners));
        for (;;) {
            Job* job = nullptr;
            {
                std::unique_lock<std::mutex> lock(mutex);
                plan_work.wait(lock, [&] {
                    return (!plan_queue.empty() && planned_ahead < lookahead) || (stopping && plan_queue.empty());
                });
                if (plan_queue.empty()) return;  // stopping, and every request has its plan
                job = plan_queue.front();
                plan_queue.pop_front();
            }
            std::unique_ptr<pm::HostPlan> host;
            std::exception_ptr error;
            try {
                const pm_request& rq = job->request;
                if (!rq.q || !rq.a || !rq.b || !rq.c) throw std::invalid_argument("null request buffer");
                host = std::make_unique<pm::HostPlan>(
                    pm::build_plan(rq.n, rq.width, rq.limbs, rq.q, pm::Tables::kDevice));
            } catch (...) {
                error = std::current_exception();
            }
            {
                std::lock_guard<std::mutex> lock(mutex);
                job->host = std::move(host);
                job->plan_error = error;
                job->planned = true;
                if (!job->taken) ++planned_ahead;
            }
            planned.notify_all();
        }
    }

    void run(pm::Lane& lane, Job& job) {
        {
            std::unique_lock<std::mutex> lock(mutex);
            planned.wait(lock, [&] { return job.planned; });
        }
        const pm_request& rq = job.request;
        job.status = guarded([&] {
            if (job.plan_error) std::rethrow_exception(job.plan_error);
            auto plan = std::make_unique<pm_plan>();
            plan->host = std::move(*job.host);
            job.host.reset();
            const size_t need = plan_footprint(plan->host);
            char* block = arena->acquire(need);
            job.admitted = Clock::now();
            // The block goes back only after the lane's streams are idle, also
            // when a step below throws.
            struct Hold {
                pm::DeviceArena* arena;
                pm::Lane* lane;
                char* block;
                ~Hold() {
                    lane->synchronize();
                    arena->release(block);
                }
            } hold{arena.get(), &lane, block};
            plan->lane = &lane;
            plan->block = block;
            plan->block_size = need;
            plan->make_device();
            plan->multiply_host(rq.a, rq.b, nullptr, rq.c);
            plan.reset();
        });
        if (job.status != PM_OK) job.error = g_last_error;
        if (job.admitted == Clock::time_point()) job.admitted = Clock::now();
        job.finished_at = Clock::now();
    }

    void worker(size_t index) {
        (void)hipSetDevice(device_id);
        pm::Lane& lane = *lanes[index];
        for (;;) {
            Job* job = nullptr;
            {
                std::unique_lock<std::mutex> lock(mutex);
                work.wait(lock, [&] { return stopping || !queue.empty(); });
                if (queue.empty()) return;  // stopping, and every request has been taken
                job = queue.front();
                queue.pop_front();
                if (job->planned)
                    --planned_ahead;
                else
                    job->taken = true;
            }
            plan_work.notify_all();  // a place in the look-ahead may have freed
            run(lane, *job);
            {
                std::lock_guard<std::mutex> lock(mutex);
                job->finished = true;
            }
            done.notify_all();
        }
    }

    void stop() {
        {
            std::lock_guard<std::mutex> lock(mutex);
            stopping = true;
        }
        work.notify_all();
        plan_work.notify_all();
        for (auto& w : workers)  // they finish every queued request (planners still run)
            if (w.joinable()) w.join();
        workers.clear();
        for (auto& p : planners)
            if (p.joinable()) p.join();
        planners.clear();
    }

    ~pm_server() { stop(); }
};

namespace {

pm_status server_create(uint64_t budget, uint32_t lanes, pm_server** out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("null output pointer");
        *out = nullptr;
        if (!lanes || lanes > 64) throw std::invalid_argument("lanes must be in 1..64");
        if (budget < (uint64_t(1) << 20)) throw std::invalid_argument("device budget below 1 MiB");
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) throw NoDevice("no HIP device");
        auto sv = std::make_unique<pm_server>();
        hip_check(hipGetDevice(&sv->device_id), "hipGetDevice");
        sv->arena = std::make_unique<pm::DeviceArena>(size_t(budget));
        const unsigned copy_threads = std::max(2u, pm::Lane::default_copy_threads() / lanes);
        for (uint32_t i = 0; i < lanes; ++i) sv->lanes.push_back(std::make_unique<pm::Lane>(copy_threads));
        sv->lookahead = 2 * size_t(lanes);
        for (uint32_t i = 0; i < lanes; ++i) sv->workers.emplace_back([s = sv.get(), i] { s->worker(i); });
        for (unsigned i = 0; i < kPlanners; ++i) sv->planners.emplace_back([s = sv.get()] { s->planner(); });
        *out = sv.release();
    });
}

}  // namespace

extern "C" {

pm_status pm_server_create(uint64_t device_budget_bytes, uint32_t lanes, pm_server** out) {
    return server_create(device_budget_bytes, lanes, out);
}

pm_status pm_server_submit(pm_server* server, const pm_request* request, uint64_t* ticket) {
    return guarded([&] {
        if (!server || !request || !ticket) throw std::invalid_argument("null argument");
        auto job = std::make_unique<Job>();
        job->request = *request;
        job->submitted = Clock::now();
        {
            std::lock_guard<std::mutex> lock(server->mutex);
            if (server->stopping) throw std::invalid_argument("server is shutting down");
            *ticket = server->next_ticket++;
            server->queue.push_back(job.get());
            server->plan_queue.push_back(job.get());
            server->jobs.emplace(*ticket, std::move(job));
        }
        server->plan_work.notify_one();
        server->work.notify_one();
    });
}

pm_status pm_server_wait(pm_server* server, uint64_t ticket, pm_request_times* times) {
    if (!server) {
        g_last_error = "null server";
        return PM_ERR_INVALID_ARGUMENT;
    }
    std::unique_ptr<Job> job;
    {
        std::unique_lock<std::mutex> lock(server->mutex);
        auto it = server->jobs.find(ticket);
        if (it == server->jobs.end()) {
            g_last_error = "unknown ticket";
            return PM_ERR_INVALID_ARGUMENT;
        }
        Job* j = it->second.get();
        if (j->claimed) {  // the first waiter frees the job when it returns
            g_last_error = "ticket is already being waited for";
            return PM_ERR_INVALID_ARGUMENT;
        }
        j->claimed = true;
        server->done.wait(lock, [&] { return j->finished; });
        job = std::move(it->second);
        server->jobs.erase(it);
    }
    if (times) {
        times->queue_ms = ms_between(job->submitted, job->admitted);
        times->service_ms = ms_between(job->admitted, job->finished_at);
        times->total_ms = ms_between(job->submitted, job->finished_at);
    }
    if (job->status != PM_OK) g_last_error = job->error;
    return job->status;
}

void pm_server_destroy(pm_server* server) { delete server; }

pm_status pm_multiply_batch(const pm_request* requests, size_t count, uint64_t device_budget_bytes, uint32_t lanes,
                            pm_status* statuses) {
    if (!requests && count) {
        g_last_error = "null requests";
        return PM_ERR_INVALID_ARGUMENT;
    }
    pm_server* server = nullptr;
    pm_status st = server_create(device_budget_bytes, lanes, &server);
    if (st != PM_OK) return st;
    std::vector<uint64_t> tickets(count, 0);
    pm_status first = PM_OK;
    std::string first_error;
    // Longest first: the whole batch is known, so the largest products start
    // while the small ones fill the lanes at the end, instead of one large
    // product running alone after the rest (a product's work grows about as
    // N * L^2, the GEMMs').
    std::vector<size_t> order(count);
    for (size_t i = 0; i < count; ++i) order[i] = i;
    auto cost = [&](size_t i) { return double(requests[i].n) * requests[i].limbs * requests[i].limbs; };
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost(x) > cost(y); });
    for (size_t i : order) {
        pm_status s = pm_server_submit(server, &requests[i], &tickets[i]);
        if (s != PM_OK) {
            tickets[i] = 0;
            if (statuses) statuses[i] = s;
            if (first == PM_OK) first = s, first_error = g_last_error;
        }
    }
    for (size_t i = 0; i < count; ++i) {
        if (!tickets[i]) continue;
        pm_status s = pm_server_wait(server, tickets[i], nullptr);
        if (statuses) statuses[i] = s;
        if (s != PM_OK && first == PM_OK) first = s, first_error = g_last_error;
    }
    pm_server_destroy(server);
    if (first != PM_OK) g_last_error = first_error;
    return first;
}

}  // extern "C"
