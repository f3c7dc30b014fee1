// fherma_init / fherma_run / fherma_free over the C ABI (include/polymul.h).
// Validation and error behaviour follow the reference's entry points
// (candidates/ordered-streaming/solve.cu:900-904, 1226-1229): an invalid point
// or a wrongly sized input throws std::runtime_error with a message; the
// harness records the case as crashed and continues.
#include "fherma.h"
#include "polymul.h"

#include <future>
#include <string>

namespace {

[[noreturn]] void fail(const std::string& where, pm_status status) {
    throw std::runtime_error(where + ": " + pm_last_error() + " (status " + std::to_string(int(status)) + ")");
}

}  // namespace

void* fherma_init(const fherma::Point& p) {
    if (p.q.data.size() != size_t(p.L))
        throw std::runtime_error("invalid point: q has " + std::to_string(p.q.data.size()) + " limbs, expected L=" +
                                 std::to_string(p.L));
    pm_plan* plan = nullptr;
    pm_status status = pm_plan_create(p.N, p.W, p.L, p.q.data.data(), &plan);
    if (status != PM_OK) fail("fherma_init", status);
    return plan;
}

fherma::Outputs fherma_run(void* state, const fherma::Inputs& in) {
    auto* plan = static_cast<pm_plan*>(state);
    if (!plan) throw std::runtime_error("fherma_run: null state");
    pm_plan_info info{};
    pm_status status = pm_plan_get_info(plan, &info);
    if (status != PM_OK) fail("fherma_run", status);
    const size_t count = size_t(info.n) * info.limbs;
    if (in.a.data.size() != count || in.b.data.size() != count)
        throw std::runtime_error("fherma_run: inputs must hold N*L = " + std::to_string(count) + " limbs");
    fherma::Outputs out;
    out.c.shape = {int64_t(info.n), int64_t(info.limbs)};
    // The returned vector is zero-filled by its constructor (N*L words, first
    // touch of fresh pages); that is built on a helper thread while the inputs
    // go in and the device works, as the reference does with its output worker
    // (candidates/ordered-streaming/output_worker.h:17-75).
    auto output = std::async(std::launch::async, [&out, count] { out.c.data.resize(count); });
    status = pm_multiply_begin(plan, in.a.data.data(), in.b.data.data());
    try {
        output.get();  // joins, and rethrows an allocation failure
    } catch (...) {
        if (status == PM_OK) pm_multiply_finish(plan, nullptr);  // leave the plan ready for the next call
        throw;
    }
    if (status != PM_OK) fail("fherma_run", status);
    status = pm_multiply_finish(plan, out.c.data.data());
    if (status != PM_OK) fail("fherma_run", status);
    return out;
}

void fherma_free(void* state) { pm_plan_destroy(static_cast<pm_plan*>(state)); }
