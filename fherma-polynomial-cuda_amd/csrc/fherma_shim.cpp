// fherma_init / fherma_run / fherma_free over the C ABI (include/polymul.h).
// Validation and error behaviour follow the reference's entry points
// (candidates/ordered-streaming/solve.cu:900-904, 1226-1229): an invalid point
// or a wrongly sized input throws std::runtime_error with a message; the
// harness records the case as crashed and continues.
#include "fherma.h"
#include "polymul.h"

#include <atomic>
#include <condition_variable>
#include <exception>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace {

[[noreturn]] void fail(const std::string& where, pm_status status) {
    throw std::runtime_error(where + ": " + pm_last_error() + " (status " + std::to_string(int(status)) + ")");
}

// Builds the returned vector (N*L words, zero-filled by its constructor: the
// first touch of its pages) on a helper thread while the inputs go in and the
// device works, as the reference does with its output worker
// (candidates/ordered-streaming/output_worker.h:17-75).  One thread per state,
// made at init and woken per call: a thread started per call costs its start-up
// and, through glibc's per-thread arenas, tends to get fresh (faulting) pages
// where a long-lived thread gets its previous output's pages back.
class OutputBuilder {
public:
    OutputBuilder() : thread_([this] { loop(); }) {}

    ~OutputBuilder() {
        {
            std::lock_guard<std::mutex> lock(mutex_);
            stop_ = true;
        }
        wake_.notify_one();
        thread_.join();
    }

    OutputBuilder(const OutputBuilder&) = delete;
    OutputBuilder& operator=(const OutputBuilder&) = delete;

    void start(std::vector<uint32_t>* target, size_t count) {
        {
            std::lock_guard<std::mutex> lock(mutex_);
            target_ = target;
            count_ = count;
            error_ = nullptr;
            busy_.store(true, std::memory_order_relaxed);
            ++generation_;
        }
        wake_.notify_one();
    }

    // Waits for the vector and rethrows an allocation failure.
    void join() {
        while (busy_.load(std::memory_order_acquire)) __builtin_ia32_pause();
        if (error_) std::rethrow_exception(std::exchange(error_, nullptr));
    }

private:
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            std::vector<uint32_t>* target;
            size_t count;
            {
                std::unique_lock<std::mutex> lock(mutex_);
                wake_.wait(lock, [&] { return stop_ || generation_ != seen; });
                if (stop_) return;
                seen = generation_;
                target = target_;
                count = count_;
            }
            try {
                target->resize(count);
            } catch (...) {
                error_ = std::current_exception();
            }
            busy_.store(false, std::memory_order_release);
        }
    }

    std::mutex mutex_;
    std::condition_variable wake_;
    uint64_t generation_ = 0;
    bool stop_ = false;
    std::vector<uint32_t>* target_ = nullptr;
    size_t count_ = 0;
    std::exception_ptr error_;
    std::atomic<bool> busy_{false};
    std::thread thread_;  // last: starts after the members it reads
};

struct FhermaState {
    pm_plan* plan = nullptr;
    OutputBuilder output;
    ~FhermaState() { pm_plan_destroy(plan); }
};

}  // namespace

void* fherma_init(const fherma::Point& p) {
    if (p.q.data.size() != size_t(p.L))
        throw std::runtime_error("invalid point: q has " + std::to_string(p.q.data.size()) + " limbs, expected L=" +
                                 std::to_string(p.L));
    auto state = std::make_unique<FhermaState>();
    pm_status status = pm_plan_create(p.N, p.W, p.L, p.q.data.data(), &state->plan);
    if (status != PM_OK) fail("fherma_init", status);
    return state.release();
}

fherma::Outputs fherma_run(void* opaque, const fherma::Inputs& in) {
    auto* state = static_cast<FhermaState*>(opaque);
    if (!state) throw std::runtime_error("fherma_run: null state");
    pm_plan* plan = state->plan;
    pm_plan_info info{};
    pm_status status = pm_plan_get_info(plan, &info);
    if (status != PM_OK) fail("fherma_run", status);
    const size_t count = size_t(info.n) * info.limbs;
    if (in.a.data.size() != count || in.b.data.size() != count)
        throw std::runtime_error("fherma_run: inputs must hold N*L = " + std::to_string(count) + " limbs");
    fherma::Outputs out;
    out.c.shape = {int64_t(info.n), int64_t(info.limbs)};
    state->output.start(&out.c.data, count);
    status = pm_multiply_begin(plan, in.a.data.data(), in.b.data.data());
    try {
        state->output.join();  // before `out` can go out of scope, also on a failed begin
    } catch (...) {
        if (status == PM_OK) pm_multiply_finish(plan, nullptr);  // leave the plan ready for the next call
        throw;
    }
    if (status != PM_OK) fail("fherma_run", status);
    status = pm_multiply_finish(plan, out.c.data.data());
    if (status != PM_OK) fail("fherma_run", status);
    return out;
}

void fherma_free(void* opaque) { delete static_cast<FhermaState*>(opaque); }
