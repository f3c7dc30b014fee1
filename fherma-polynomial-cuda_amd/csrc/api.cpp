// C ABI of the gfx950 negacyclic multiplier (include/polymul.h): plan
// lifetime, device memory, the per-product stage sequence and parity taps.
// The stage sequence follows fherma_run's non-pipelined path
// (candidates/ordered-streaming/solve.cu:1230-1270): H2D, conversion, NTT
// stages, CRT GEMM, reconstruction, D2H, with integer-only reconstruction for
// every q instead of the reference's FP64 fallbacks (solve.cu:718-895).
#include "polymul.h"

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>

#include "state.hpp"

thread_local std::string g_last_error;

namespace {

pm_status create(uint32_t n, uint32_t width, uint32_t limbs, const uint32_t* q, pm_plan** out, bool device,
                 uint32_t min_block = 1) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("null output pointer");
        *out = nullptr;
        auto plan = std::make_unique<pm_plan>();
        plan->host = pm::build_plan(n, width, limbs, q, device ? pm::Tables::kDevice : pm::Tables::kHost, min_block);
        if (device) plan->make_device();
        *out = plan.release();
    });
}

template <class T>
void copy_table(const std::vector<T>& v, void* out, size_t* bytes) {
    size_t need = v.size() * sizeof(T);
    if (out) {
        if (*bytes < need) throw std::invalid_argument("table buffer too small");
        std::memcpy(out, v.data(), need);
    }
    *bytes = need;
}

}  // namespace

extern "C" {

pm_status pm_plan_create(uint32_t n, uint32_t width, uint32_t limbs, const uint32_t* q, pm_plan** out) {
    return create(n, width, limbs, q, out, true);
}

pm_status pm_plan_create_host(uint32_t n, uint32_t width, uint32_t limbs, const uint32_t* q, pm_plan** out) {
    return create(n, width, limbs, q, out, false);
}

pm_status pm_plan_create_split(uint32_t n, uint32_t width, uint32_t limbs, const uint32_t* q, uint32_t min_block,
                               int device, pm_plan** out) {
    return create(n, width, limbs, q, out, device != 0, min_block);
}

void pm_plan_destroy(pm_plan* plan) { delete plan; }

const char* pm_last_error(void) { return g_last_error.c_str(); }

pm_status pm_multiply(pm_plan* plan, const uint32_t* a, const uint32_t* b, uint32_t* c) {
    return guarded([&] {
        require_device(plan);
        if (!a || !b || !c) throw std::invalid_argument("null operand");
        hip_check(hipSetDevice(plan->device_id), "hipSetDevice");
        plan->multiply_host(a, b, nullptr, c);
    });
}

pm_status pm_multiply_begin(pm_plan* plan, const uint32_t* a, const uint32_t* b) {
    return guarded([&] {
        require_device(plan);
        if (!a || !b) throw std::invalid_argument("null operand");
        hip_check(hipSetDevice(plan->device_id), "hipSetDevice");
        plan->multiply_begin(a, b, nullptr);
    });
}

pm_status pm_multiply_finish(pm_plan* plan, uint32_t* c) {
    return guarded([&] {
        require_device(plan);
        hip_check(hipSetDevice(plan->device_id), "hipSetDevice");
        plan->multiply_finish(c);
    });
}

pm_status pm_multiply_device(pm_plan* plan, const uint32_t* a, const uint32_t* b, uint32_t* c, void* stream) {
    return guarded([&] {
        require_device(plan);
        require_idle(plan);
        if (!a || !b || !c) throw std::invalid_argument("null operand");
        hip_check(hipSetDevice(plan->device_id), "hipSetDevice");
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : plan->lane->stream;
        plan->run(a, b, c, PM_STAGE_OUTPUT, st);
        hip_check(hipStreamSynchronize(st), "multiply");
        plan->collect_times();
    });
}

namespace {

// Per device: the buffer of pm_multiply_device_batch's items and work lists,
// kept for the process (a few tens of KB per hundred products).
struct BatchBuffer {
    std::mutex mutex;
    void* ptr = nullptr;
    size_t bytes = 0;
};

BatchBuffer& batch_buffer(int device) {
    static std::mutex guard;
    static std::map<int, std::unique_ptr<BatchBuffer>> buffers;
    std::lock_guard<std::mutex> lock(guard);
    std::unique_ptr<BatchBuffer>& b = buffers[device];
    if (!b) b = std::make_unique<BatchBuffer>();
    return *b;
}

}  // namespace

pm_status pm_multiply_device_batch(pm_plan* const* plans, const uint32_t* const* a, const uint32_t* const* b,
                                   uint32_t* const* c, size_t count, void* stream) {
    return guarded([&] {
        if (count == 0) return;
        if (!plans || !a || !b || !c) throw std::invalid_argument("null argument");
        std::vector<const pm_plan*> seen;
        seen.reserve(count);
        int device = -1;
        for (size_t i = 0; i < count; ++i) {
            if (!plans[i]) throw std::invalid_argument("null plan");
            require_device(plans[i]);
            require_idle(plans[i]);
            if (!a[i] || !b[i] || !c[i]) throw std::invalid_argument("null operand");
            if (device < 0) device = plans[i]->device_id;
            if (plans[i]->device_id != device) throw std::invalid_argument("plans of one batch on different devices");
            seen.push_back(plans[i]);
        }
        std::sort(seen.begin(), seen.end());
        if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
            throw std::invalid_argument("a plan appears twice in one batch (its working buffers hold one product)");
        hip_check(hipSetDevice(device), "hipSetDevice");
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : plans[0]->lane->stream;
        // A failure after some launches are queued waits for them before it is
        // reported: they use the plans' buffers and the batch buffer.
        try {
            std::vector<pm::GroupItem> items;
            for (size_t i = 0; i < count; ++i) {
                const pm::Shape& s = plans[i]->host.shape;
                if (pm::groupable(s))
                    items.push_back(pm::group_item(s, plans[i]->tables, plans[i]->work, a[i], b[i], c[i]));
                else
                    plans[i]->run(a[i], b[i], c[i], PM_STAGE_OUTPUT, st);  // on its own, ahead of the grouped launches
            }
            if (!items.empty()) {
                // The items and the launches' work lists go into the device's
                // batch buffer, held (and grown) across calls; a second batch on
                // the same device waits here until the first has finished with it.
                BatchBuffer& buf = batch_buffer(device);
                std::lock_guard<std::mutex> hold(buf.mutex);
                try {  // the stream is idle again before the buffer is let go
                    const size_t need = pm::group_bytes(items.size());
                    if (buf.bytes < need) {
                        if (buf.ptr) hip_check(hipFree(buf.ptr), "hipFree");
                        buf.ptr = nullptr;
                        buf.bytes = 0;
                        hip_check(hipMalloc(&buf.ptr, need), "hipMalloc");
                        buf.bytes = need;
                    }
                    pm::launch_group(items, buf.ptr, st);
                    hip_check(hipStreamSynchronize(st), "multiply batch");
                } catch (...) {
                    (void)hipStreamSynchronize(st);
                    throw;
                }
                return;
            }
        } catch (...) {
            (void)hipStreamSynchronize(st);
            throw;
        }
        hip_check(hipStreamSynchronize(st), "multiply batch");
    });
}

pm_status pm_multiply_device_scratch(pm_plan* plan, const uint32_t* a, const uint32_t* b, uint32_t* c,
                                     void* stream) {
    return guarded([&] {
        require_device(plan);
        require_idle(plan);
        if (!a || !b || !c) throw std::invalid_argument("null operand");
        hip_check(hipSetDevice(plan->device_id), "hipSetDevice");
        const pm::Shape& s = plan->host.shape;
        pm::DeviceWork w;
        if (!s.trivial) {
            const size_t need = pm::work_bytes(s);
            if (need > pm::kScratchBytes)
                throw std::invalid_argument("pm_multiply_device_scratch: the product needs " + std::to_string(need) +
                                            " bytes of workspace, more than the " +
                                            std::to_string(pm::kScratchBytes) + "-byte static scratch");
            std::string why;
            void* base = pm::scratch_address(plan->device_id, &why);
            if (!base) throw DeviceError("pm_multiply_device_scratch: " + why);
            pm::carve_work(s, static_cast<uint8_t*>(base), &w);
        }
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : plan->lane->stream;
        plan->run_with(w, a, b, c, PM_STAGE_OUTPUT, st);
        hip_check(hipStreamSynchronize(st), "multiply");
        plan->collect_times();
    });
}

pm_status pm_operand_create(pm_plan* plan, const uint32_t* b, pm_operand** out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("null output pointer");
        *out = nullptr;
        require_device(plan);
        require_idle(plan);
        if (!b) throw std::invalid_argument("null operand");
        hip_check(hipSetDevice(plan->device_id), "hipSetDevice");
        const pm::Shape& s = plan->host.shape;
        auto op = std::make_unique<pm_operand>();
        op->plan = plan;
        if (s.trivial) {
            *out = op.release();
            return;
        }
        const size_t plane_words = size_t(s.r) * s.n_pad;
        void* mem = nullptr;
        hipError_t e = hipMalloc(&mem, plane_words * sizeof(uint32_t));
        if (e == hipErrorOutOfMemory) throw std::bad_alloc();
        hip_check(e, "hipMalloc");
        op->planes = static_cast<uint32_t*>(mem);
        hipStream_t st = plan->lane->stream;
        const size_t bytes = size_t(s.n) * s.limbs * sizeof(uint32_t);
        const pm::TileSet all = pm::all_tiles(s);
        hip_check(hipMemcpyAsync(plan->io_b, b, bytes, hipMemcpyHostToDevice, st), "copy b to device");
        pm::launch_pack(s, nullptr, plan->io_b, plan->work.packed, all, 1, 1, st);
        pm::launch_conversion(s, plan->tables, plan->work.packed, plan->work.planes, all, 1, 1, st);
        pm::launch_ntt_forward_columns(s, plan->tables, plan->work.planes, 1, 1, st);
        hip_check(hipMemcpyAsync(op->planes, plan->b_planes(), plane_words * sizeof(uint32_t),
                                 hipMemcpyDeviceToDevice, st),
                  "keep b residues");
        hip_check(hipStreamSynchronize(st), "operand");
        *out = op.release();
    });
}

pm_status pm_multiply_operand(pm_plan* plan, const uint32_t* a, const pm_operand* b, uint32_t* c) {
    return guarded([&] {
        require_device(plan);
        if (!a || !b || !c) throw std::invalid_argument("null operand");
        if (b->plan != plan) throw std::invalid_argument("operand belongs to another plan");
        hip_check(hipSetDevice(plan->device_id), "hipSetDevice");
        plan->multiply_host(a, nullptr, b->planes, c);
    });
}

pm_status pm_multiply_operand_device(pm_plan* plan, const uint32_t* a, const pm_operand* b, uint32_t* c,
                                     void* stream) {
    return guarded([&] {
        require_device(plan);
        if (!a || !b || !c) throw std::invalid_argument("null operand");
        if (b->plan != plan) throw std::invalid_argument("operand belongs to another plan");
        require_idle(plan);
        hip_check(hipSetDevice(plan->device_id), "hipSetDevice");
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : plan->lane->stream;
        plan->run(a, nullptr, c, PM_STAGE_OUTPUT, st, b->planes);
        hip_check(hipStreamSynchronize(st), "multiply");
        plan->collect_times();
    });
}

void pm_operand_destroy(pm_operand* operand) { delete operand; }

pm_status pm_plan_get_info(const pm_plan* plan, pm_plan_info* info) {
    return guarded([&] {
        if (!plan || !info) throw std::invalid_argument("null argument");
        const pm::Shape& s = plan->host.shape;
        info->n = s.n;
        info->width = s.width;
        info->limbs = s.limbs;
        info->primes = s.r;
        info->block = s.n1;
        info->transform = s.n2;
        info->row_width = s.row_width;
        info->reference_walk = s.reference_walk ? 1 : 0;
        info->qbits = s.qbits;
        info->trivial = s.trivial ? 1 : 0;
        info->device_bytes = plan->device_bytes;
    });
}

pm_status pm_plan_get_primes(const pm_plan* plan, uint32_t* primes, size_t capacity) {
    return guarded([&] {
        if (!plan || (!primes && capacity)) throw std::invalid_argument("null argument");
        const auto& v = plan->host.primes;
        for (size_t i = 0; i < v.size() && i < capacity; ++i) primes[i] = v[i];
    });
}

pm_status pm_plan_export_table(const pm_plan* plan, int table, void* out, size_t* bytes) {
    return guarded([&] {
        if (!plan || !bytes) throw std::invalid_argument("null argument");
        // device-built tables are fetched (and untiled) on first export
        const_cast<pm_plan*>(plan)->fetch_device_tables();
        const pm::HostPlan& h = plan->host;
        switch (table) {
            case PM_TABLE_PRIME_CONSTS: copy_table(h.consts, out, bytes); break;
            case PM_TABLE_CONV_WEIGHTS: copy_table(h.wconv, out, bytes); break;
            case PM_TABLE_TWIDDLES_FWD: copy_table(h.tw_fwd, out, bytes); break;
            case PM_TABLE_TWIDDLES_INV: copy_table(h.tw_inv, out, bytes); break;
            case PM_TABLE_CRT_WEIGHTS: copy_table(h.wcrt, out, bytes); break;
            case PM_TABLE_M_MOD_Q: copy_table(h.m_mod_q, out, bytes); break;
            case PM_TABLE_NEG_P_MOD_Q: copy_table(h.neg_p_mod_q, out, bytes); break;
            case PM_TABLE_CORR_T: copy_table(h.corr_t, out, bytes); break;
            case PM_TABLE_CORR_E: copy_table(h.corr_e, out, bytes); break;
            case PM_TABLE_RECIPROCALS: copy_table(h.reciprocal, out, bytes); break;
            case PM_TABLE_BETA: copy_table(h.beta, out, bytes); break;
            case PM_TABLE_SHAPE: {
                const pm::Shape& s = h.shape;
                std::vector<uint32_t> w = {s.n,         s.width,     s.limbs,         s.n1,
                                           s.n2,        s.log_n2,    s.r,             s.r_pad,
                                           s.conv_k,    s.conv_rows, s.crt_words,     s.crt_words_pad,
                                           s.crt_rows,  s.crt_k,     s.t_words,       s.row_width,
                                           s.rows,      s.qbits,     s.s_shift,       s.k_shift,
                                           uint32_t(s.mu), uint32_t(s.mu >> 32), s.trivial ? 1u : 0u,
                                           s.reference_walk ? 1u : 0u};
                copy_table(w, out, bytes);
                break;
            }
            default: throw std::invalid_argument("unknown table");
        }
    });
}

pm_status pm_debug_stage(pm_plan* plan, const uint32_t* a, const uint32_t* b, int stage, void* out, size_t* bytes) {
    return guarded([&] {
        if (!plan || !bytes) throw std::invalid_argument("null argument");
        const pm::Shape& s = plan->host.shape;
        size_t need = 0;
        switch (stage) {
            case PM_STAGE_RESIDUES: need = size_t(2) * s.r * s.n * 4; break;
            case PM_STAGE_PRODUCT: need = size_t(s.r) * s.n * 4; break;
            case PM_STAGE_WORDS: need = size_t(s.crt_words) * s.n * 8; break;
            case PM_STAGE_OUTPUT: need = size_t(s.n) * s.limbs * 4; break;
            default: throw std::invalid_argument("unknown stage");
        }
        if (!out) {
            *bytes = need;
            return;
        }
        if (*bytes < need) throw std::invalid_argument("stage buffer too small");
        require_device(plan);
        require_idle(plan);
        if (!a || !b) throw std::invalid_argument("null operand");
        if (s.trivial && stage != PM_STAGE_OUTPUT) throw std::invalid_argument("q == 1 has no intermediate stages");
        hip_check(hipSetDevice(plan->device_id), "hipSetDevice");
        const size_t nl = size_t(s.n) * s.limbs * 4;
        hipStream_t st = plan->lane->stream;
        hip_check(hipMemcpyAsync(plan->io_a, a, nl, hipMemcpyHostToDevice, st), "copy a");
        hip_check(hipMemcpyAsync(plan->io_b, b, nl, hipMemcpyHostToDevice, st), "copy b");
        plan->run(plan->io_a, plan->io_b, plan->io_c, stage, st);
        if (stage == PM_STAGE_OUTPUT) {
            hip_check(hipMemcpyAsync(out, plan->io_c, need, hipMemcpyDeviceToHost, st), "copy stage");
        } else if (stage == PM_STAGE_PRODUCT) {
            // v_k lives in the CRT operand tiles: [n_pad/256][r_pad/16][256][16 x u32], XOR 0x80808080
            const size_t tile_words = size_t(s.r_pad) * s.n_pad;
            std::vector<uint32_t> tiles(tile_words);
            hip_check(hipMemcpyAsync(tiles.data(), plan->work.rtiles, tile_words * 4, hipMemcpyDeviceToHost, st),
                      "copy stage");
            hip_check(hipStreamSynchronize(st), "debug stage");
            uint32_t* v = static_cast<uint32_t*>(out);
            const size_t ksteps = s.r_pad / 16;
            for (size_t k = 0; k < s.r; ++k)
                for (size_t i = 0; i < s.n; ++i)
                    v[k * s.n + i] = tiles[(((i >> 8) * ksteps + k / 16) * 256 + (i & 255)) * 16 + k % 16] ^ 0x80808080u;
        } else {
            // device rows have stride n_pad; the tap returns rows of exactly N
            const size_t elem = stage == PM_STAGE_WORDS ? 8 : 4;
            if (stage == PM_STAGE_WORDS && pm::words_by_coefficient(s)) {
                // coefficient-major on the device: transpose into the tap's [crt_words][N]
                const size_t ws = pm::word_stride(s);
                std::vector<int64_t> sums(size_t(s.n) * ws);
                hip_check(hipMemcpyAsync(sums.data(), plan->work.words, sums.size() * 8, hipMemcpyDeviceToHost, st),
                          "copy stage");
                hip_check(hipStreamSynchronize(st), "debug stage");
                int64_t* w = static_cast<int64_t*>(out);
                for (size_t j = 0; j < s.crt_words; ++j)
                    for (size_t i = 0; i < s.n; ++i) w[j * s.n + i] = sums[i * ws + j];
            } else if (stage == PM_STAGE_WORDS) {
                hip_check(hipMemcpy2DAsync(out, s.n * elem, plan->work.words, size_t(s.n_pad) * elem, s.n * elem,
                                           s.crt_words, hipMemcpyDeviceToHost, st),
                          "copy stage");
            } else {  // residues of a then b; the b plane starts r_pad rows in
                for (size_t op = 0; op < 2; ++op)
                    hip_check(hipMemcpy2DAsync(static_cast<uint8_t*>(out) + op * s.r * s.n * elem, s.n * elem,
                                               plan->work.planes + op * s.r_pad * size_t(s.n_pad),
                                               size_t(s.n_pad) * elem, s.n * elem, s.r, hipMemcpyDeviceToHost, st),
                              "copy stage");
            }
        }
        hip_check(hipStreamSynchronize(st), "debug stage");
        *bytes = need;
    });
}

pm_status pm_set_profiling(pm_plan* plan, int enabled) {
    return guarded([&] {
        if (!plan) throw std::invalid_argument("null plan");
        plan->profiling = enabled != 0;
    });
}

pm_status pm_get_stage_times(const pm_plan* plan, float* ms, int count) {
    return guarded([&] {
        if (!plan || (!ms && count > 0)) throw std::invalid_argument("null argument");
        for (int i = 0; i < count && i < kStages; ++i) ms[i] = plan->stage_ms[i];
    });
}

}  // extern "C"
