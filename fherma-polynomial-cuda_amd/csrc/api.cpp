// C ABI of the gfx950 negacyclic multiplier (include/polymul.h): plan
// lifetime, device memory, the per-product stage sequence and parity taps.
// The stage sequence follows fherma_run's non-pipelined path
// (candidates/ordered-streaming/solve.cu:1230-1270): H2D, conversion, NTT
// stages, CRT GEMM, reconstruction, D2H, with integer-only reconstruction for
// every q instead of the reference's FP64 fallbacks (solve.cu:718-895).
#include "polymul.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <algorithm>
#include <memory>
#include <thread>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "copy_pool.hpp"
#include "kernels.hpp"
#include "plan.hpp"

namespace {

thread_local std::string g_last_error;

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct NoDevice : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* where) {
    if (e != hipSuccess) throw DeviceError(std::string(where) + ": " + hipGetErrorString(e));
}

// Events bracket the 7 stage launches; times are the 7 intervals plus the total.
constexpr int kEvents = 8;
constexpr int kStages = 8;
// The host-vector path moves operands in and results out in up to this many
// coefficient chunks, overlapping copies, transfers and the per-coefficient stages.
constexpr int kMaxChunks = 8;

}  // namespace

struct pm_plan {
    pm::HostPlan host;
    bool device = false;
    int device_id = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocations;
    pm::DeviceTables tables;
    pm::DeviceWork work;
    uint32_t* io_a = nullptr;  // device copies for the host-vector entry point
    uint32_t* io_b = nullptr;
    uint32_t* io_c = nullptr;
    // host staging for pm_multiply (created on first use)
    uint32_t* pin_in = nullptr;   // [2][N][L] pinned
    uint32_t* pin_out = nullptr;  // [N][L] pinned
    hipStream_t copy_in = nullptr, copy_out = nullptr;
    hipEvent_t in_ready[kMaxChunks] = {}, out_ready[kMaxChunks] = {}, reconstructed[kMaxChunks] = {};
    std::unique_ptr<pm::CopyPool> pool;
    uint64_t device_bytes = 0;
    uint32_t* dev_beta = nullptr;  // device-built beta (Tables::kDevice), for table export
    bool profiling = false;
    hipEvent_t events[kEvents] = {};
    float stage_ms[kStages] = {};

    template <class T>
    T* alloc(size_t count) {
        void* p = nullptr;
        size_t bytes = count * sizeof(T);
        if (bytes == 0) bytes = 256;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            if (e == hipErrorOutOfMemory) throw std::bad_alloc();
            hip_check(e, "hipMalloc");
        }
        allocations.push_back(p);
        device_bytes += bytes;
        return static_cast<T*>(p);
    }

    template <class T>
    T* upload(const std::vector<T>& v) {
        T* p = alloc<T>(v.size());
        if (!v.empty()) hip_check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload plan");
        return p;
    }

    ~pm_plan() {
        if (device) {
            hipSetDevice(device_id);
            for (hipStream_t st : {stream, copy_in, copy_out})
                if (st) hipStreamSynchronize(st);
            for (hipEvent_t& e : events)
                if (e) hipEventDestroy(e);
            for (int i = 0; i < kMaxChunks; ++i)
                for (hipEvent_t e : {in_ready[i], out_ready[i], reconstructed[i]})
                    if (e) hipEventDestroy(e);
            for (void* p : allocations) hipFree(p);
            if (pin_in) hipHostFree(pin_in);
            if (pin_out) hipHostFree(pin_out);
            for (hipStream_t st : {stream, copy_in, copy_out})
                if (st) hipStreamDestroy(st);
        }
    }

    void make_staging() {
        if (pin_in) return;
        const size_t bytes = size_t(host.shape.n) * host.shape.limbs * sizeof(uint32_t);
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&pin_in), 2 * bytes, hipHostMallocDefault),
                  "pin input staging");
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&pin_out), bytes, hipHostMallocDefault),
                  "pin output staging");
        hip_check(hipStreamCreateWithFlags(&copy_in, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipStreamCreateWithFlags(&copy_out, hipStreamNonBlocking), "hipStreamCreate");
        for (int i = 0; i < kMaxChunks; ++i) {
            hip_check(hipEventCreateWithFlags(&in_ready[i], hipEventDisableTiming), "hipEventCreate");
            hip_check(hipEventCreateWithFlags(&out_ready[i], hipEventDisableTiming), "hipEventCreate");
            hip_check(hipEventCreateWithFlags(&reconstructed[i], hipEventDisableTiming), "hipEventCreate");
        }
        unsigned hw = std::thread::hardware_concurrency();
        pool = std::make_unique<pm::CopyPool>(hw ? std::min(16u, hw) : 4u);
    }

    // Host vectors in, host vector out.  Chunk c (256-coefficient tiles
    // [t0, t0+tc)) is copied into pinned memory by the pool, sent on copy_in,
    // then packed and converted on the compute stream while the next chunk is
    // still being copied; after the transforms and the CRT GEMM, results are
    // reconstructed chunk by chunk and each chunk's D2H and host copy-out
    // overlap the next chunk's reconstruction.
    // b == nullptr: the b residues come from `bplanes` (a resident operand).
    void multiply_host(const uint32_t* a, const uint32_t* b, const uint32_t* bplanes_in, uint32_t* c) {
        const pm::Shape& s = host.shape;
        const uint32_t nops = b ? 2 : 1;
        const uint32_t* bplanes = b ? b_planes() : bplanes_in;
        const size_t L = s.limbs, row_bytes = L * sizeof(uint32_t);
        if (s.trivial) {
            std::memset(c, 0, size_t(s.n) * row_bytes);
            return;
        }
        make_staging();
        const uint32_t ntiles = s.n_pad / 256;
        const uint32_t chunks = ntiles < kMaxChunks ? ntiles : kMaxChunks;
        const uint32_t per = (ntiles + chunks - 1) / chunks;
        uint32_t* pin_a = pin_in;
        uint32_t* pin_b = pin_in + size_t(s.n) * L;
        auto rows_of = [&](uint32_t ch, size_t& r0, size_t& r1) {
            r0 = std::min<size_t>(size_t(ch) * per * 256, s.n);
            r1 = std::min<size_t>(size_t(ch + 1) * per * 256, s.n);
        };
        for (uint32_t ch = 0; ch < chunks; ++ch) {
            size_t r0, r1;
            rows_of(ch, r0, r1);
            if (r1 > r0) {
                const size_t off = r0 * L, bytes = (r1 - r0) * row_bytes;
                pool->copy(pin_a + off, a + off, bytes);
                hip_check(hipMemcpyAsync(io_a + off, pin_a + off, bytes, hipMemcpyHostToDevice, copy_in), "H2D a");
                if (b) {
                    pool->copy(pin_b + off, b + off, bytes);
                    hip_check(hipMemcpyAsync(io_b + off, pin_b + off, bytes, hipMemcpyHostToDevice, copy_in),
                              "H2D b");
                }
            }
            hip_check(hipEventRecord(in_ready[ch], copy_in), "hipEventRecord");
            hip_check(hipStreamWaitEvent(stream, in_ready[ch], 0), "hipStreamWaitEvent");
            const uint32_t t0 = ch * per, tc = std::min(per, ntiles - t0);
            if (tc) {
                pm::launch_pack(s, io_a, io_b, work.packed, t0, tc, 0, nops, stream);
                pm::launch_conversion(s, tables, work.packed, work.planes, t0, tc, 0, nops, stream);
            }
        }
        run_middle(stream, nops, bplanes);
        for (uint32_t ch = 0; ch < chunks; ++ch) {
            size_t r0, r1;
            rows_of(ch, r0, r1);
            pm::launch_reconstruct(s, tables, work.words, work.t_words, io_c, uint32_t(r0), uint32_t(r1 - r0),
                                   stream);
            hip_check(hipEventRecord(reconstructed[ch], stream), "hipEventRecord");
            hip_check(hipStreamWaitEvent(copy_out, reconstructed[ch], 0), "hipStreamWaitEvent");
            if (r1 > r0)
                hip_check(hipMemcpyAsync(pin_out + r0 * L, io_c + r0 * L, (r1 - r0) * row_bytes,
                                         hipMemcpyDeviceToHost, copy_out),
                          "D2H c");
            hip_check(hipEventRecord(out_ready[ch], copy_out), "hipEventRecord");
        }
        for (uint32_t ch = 0; ch < chunks; ++ch) {
            size_t r0, r1;
            rows_of(ch, r0, r1);
            hip_check(hipEventSynchronize(out_ready[ch]), "D2H c");
            if (r1 > r0) pool->copy(c + r0 * L, pin_out + r0 * L, (r1 - r0) * row_bytes);
        }
        hip_check(hipStreamSynchronize(stream), "multiply");
    }

    void make_device() {
        hip_check(hipGetDevice(&device_id), "hipGetDevice");
        hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        device = true;
        for (hipEvent_t& e : events) hip_check(hipEventCreate(&e), "hipEventCreate");
        const pm::Shape& s = host.shape;
        const size_t nl = size_t(s.n) * s.limbs;
        io_a = alloc<uint32_t>(nl);
        io_b = alloc<uint32_t>(nl);
        io_c = alloc<uint32_t>(nl);
        if (s.trivial) return;
        pm::configure_kernels();
        if (host.tables == pm::Tables::kDevice) {
            build_device_tables();
        } else {
            tables.consts = upload(host.consts);
            // both operands' weight matrices, each in 256x64 tiles
            const size_t one = size_t(s.conv_rows) * s.conv_k;
            std::vector<int8_t> tiled;
            tiled.reserve(2 * one);
            for (int op = 0; op < 2; ++op) {
                std::vector<int8_t> part(host.wconv.begin() + op * one, host.wconv.begin() + (op + 1) * one);
                std::vector<int8_t> t = pm::tile_256x64(part, s.conv_rows, s.conv_k);
                tiled.insert(tiled.end(), t.begin(), t.end());
            }
            tables.wconv = upload(tiled);
            tables.tw_fwd = reinterpret_cast<const uint2*>(upload(host.tw_fwd));
            tables.tw_inv = reinterpret_cast<const uint2*>(upload(host.tw_inv));
            tables.wcrt = upload(pm::tile_256x64(host.wcrt, s.crt_rows, s.crt_k));
        }
        tables.q = upload(host.q);
        tables.neg_p = upload(host.neg_p_mod_q);
        tables.corr_t = upload(host.corr_t);
        tables.corr_e = upload(host.corr_e);
        work.packed = alloc<int8_t>(pm::packed_operand_bytes(s));
        // planes: [2][r_pad][n_pad].  After the row stages the b half is dead and
        // receives the CRT operand tiles of v_k (exactly r_pad*n_pad words);
        // after the CRT GEMM the a half is dead and holds T ([L+2][n_pad]).
        const size_t half = size_t(s.r_pad > s.t_words ? s.r_pad : s.t_words) * s.n_pad;
        work.planes = alloc<uint32_t>(2 * half);
        work.rtiles = reinterpret_cast<int8_t*>(work.planes + half);
        work.t_words = work.planes;
        work.words = alloc<int64_t>(size_t(s.crt_words) * s.n_pad);
        // pinned staging, streams and copy workers of the host entry points are
        // set up here, with the rest of the plan, not inside the first product
        make_staging();
    }

    // Prime-set tables and CRT weights built by kernels from the O(r) and O(rL)
    // host results (psi, reciprocals, M_k): no O(r^2) or O(r*N) host loops and no
    // upload of the large tables.
    void build_device_tables() {
        const pm::Shape& s = host.shape;
        pm::TableBuild tb;
        tb.consts = upload(host.consts);
        tb.psi = upload(host.psi);
        tb.rcp = upload(host.reciprocal);
        tb.m_mod_q = upload(host.m_mod_q);
        tb.beta = dev_beta = alloc<uint32_t>(s.r);
        tb.wconv = alloc<int8_t>(size_t(2) * s.conv_rows * s.conv_k);
        tb.tw_fwd = alloc<uint32_t>(size_t(s.r) * s.n2 * 2);
        tb.tw_inv = alloc<uint32_t>(size_t(s.r) * s.n2 * 2);
        tb.wcrt = alloc<int8_t>(size_t(s.crt_rows) * s.crt_k);
        pm::launch_build_tables(s, tb, stream);
        hip_check(hipStreamSynchronize(stream), "build tables");
        tables.consts = tb.consts;
        tables.wconv = tb.wconv;
        tables.tw_fwd = reinterpret_cast<const uint2*>(tb.tw_fwd);
        tables.tw_inv = reinterpret_cast<const uint2*>(tb.tw_inv);
        tables.wcrt = tb.wcrt;
    }

    // Host copies of device-built tables, in the host layouts of plan.cpp (for
    // table export); fetched once.
    void fetch_device_tables() {
        const pm::Shape& s = host.shape;
        if (host.tables != pm::Tables::kDevice || !device || s.trivial || !host.wconv.empty()) return;
        hip_check(hipSetDevice(device_id), "hipSetDevice");
        auto get = [&](auto& vec, const void* src, size_t count) {
            vec.resize(count);
            hip_check(hipMemcpy(vec.data(), src, count * sizeof(vec[0]), hipMemcpyDeviceToHost), "fetch tables");
        };
        get(host.consts, tables.consts, s.r);
        get(host.beta, dev_beta, s.r);
        get(host.tw_fwd, tables.tw_fwd, size_t(s.r) * s.n2 * 2);
        get(host.tw_inv, tables.tw_inv, size_t(s.r) * s.n2 * 2);
        std::vector<int8_t> tiled;
        const size_t one = size_t(s.conv_rows) * s.conv_k;
        get(tiled, tables.wconv, 2 * one);
        host.wconv.resize(2 * one);
        for (int op = 0; op < 2; ++op) {
            std::vector<int8_t> part(tiled.begin() + op * one, tiled.begin() + (op + 1) * one);
            std::vector<int8_t> flat = pm::untile_256x64(part, s.conv_rows, s.conv_k);
            std::copy(flat.begin(), flat.end(), host.wconv.begin() + op * one);
        }
        get(tiled, tables.wcrt, size_t(s.crt_rows) * s.crt_k);
        host.wcrt = pm::untile_256x64(tiled, s.crt_rows, s.crt_k);
    }

    // The b residues inside the work planes ([r_pad][n_pad] after a's).
    uint32_t* b_planes() { return work.planes + size_t(host.shape.r_pad) * host.shape.n_pad; }

    // Runs stages up to and including `last` (PM_STAGE_*) on device operands;
    // b == nullptr takes the b residues from `bplanes_in` (a resident operand).
    void run(const uint32_t* a, const uint32_t* b, uint32_t* c, int last, hipStream_t st,
             const uint32_t* bplanes_in = nullptr) {
        const pm::Shape& s = host.shape;
        if (s.trivial) {
            pm::launch_zero(c, size_t(s.n) * s.limbs, st);
            return;
        }
        auto mark = [&](int i) {
            if (profiling) hip_check(hipEventRecord(events[i], st), "hipEventRecord");
        };
        const uint32_t ntiles = s.n_pad / 256;
        const uint32_t nops = b ? 2 : 1;
        const uint32_t* bplanes = b ? b_planes() : bplanes_in;
        mark(0);
        pm::launch_pack(s, a, b, work.packed, 0, ntiles, 0, nops, st);
        mark(1);
        pm::launch_conversion(s, tables, work.packed, work.planes, 0, ntiles, 0, nops, st);
        mark(2);
        if (last == PM_STAGE_RESIDUES) return;
        pm::launch_ntt_forward_columns(s, tables, work.planes, 0, nops, st);
        mark(3);
        pm::launch_ntt_rows_product(s, tables, work.planes, bplanes, st);
        mark(4);
        pm::launch_ntt_inverse_columns(s, tables, work.planes, work.rtiles, st);
        mark(5);
        if (last == PM_STAGE_PRODUCT) return;
        pm::launch_crt(s, tables, work.rtiles, work.words, st);
        mark(6);
        if (last == PM_STAGE_WORDS) return;
        pm::launch_reconstruct(s, tables, work.words, work.t_words, c, 0, s.n, st);
        mark(7);
    }

    // The stages between the per-coefficient ends: transforms and the CRT GEMM.
    void run_middle(hipStream_t st, uint32_t nops, const uint32_t* bplanes) {
        const pm::Shape& s = host.shape;
        pm::launch_ntt_forward_columns(s, tables, work.planes, 0, nops, st);
        pm::launch_ntt_rows_product(s, tables, work.planes, bplanes, st);
        pm::launch_ntt_inverse_columns(s, tables, work.planes, work.rtiles, st);
        pm::launch_crt(s, tables, work.rtiles, work.words, st);
    }

    void collect_times() {
        if (!profiling || host.shape.trivial) return;
        for (int i = 0; i < kEvents - 1; ++i)
            hip_check(hipEventElapsedTime(&stage_ms[i], events[i], events[i + 1]), "hipEventElapsedTime");
        hip_check(hipEventElapsedTime(&stage_ms[kEvents - 1], events[0], events[kEvents - 1]), "hipEventElapsedTime");
    }
};

// A resident factor: b's residues after conversion and the row-spanning
// forward stages, [r][n_pad] u32 on the plan's device.  Valid only with the
// plan that made it (b's weights carry that plan's scale beta_k).
struct pm_operand {
    pm_plan* plan = nullptr;
    uint32_t* planes = nullptr;
    ~pm_operand() {
        if (planes) {
            hipSetDevice(plan->device_id);
            hipFree(planes);
        }
    }
};

namespace {

template <class F>
pm_status guarded(F&& body) {
    try {
        body();
        return PM_OK;
    } catch (const NoDevice& e) {
        g_last_error = e.what();
        return PM_ERR_NO_DEVICE;
    } catch (const DeviceError& e) {
        g_last_error = e.what();
        return PM_ERR_DEVICE;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return PM_ERR_INVALID_ARGUMENT;
    } catch (const std::bad_alloc&) {
        g_last_error = "out of memory";
        return PM_ERR_OUT_OF_MEMORY;
    } catch (const std::runtime_error& e) {
        g_last_error = e.what();
        return std::string(e.what()).find("insufficient independent NTT fields") != std::string::npos
                   ? PM_ERR_INSUFFICIENT_FIELDS
                   : PM_ERR_INTERNAL;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return PM_ERR_INTERNAL;
    }
}

void require_device(const pm_plan* plan) {
    if (!plan) throw std::invalid_argument("null plan");
    if (!plan->device) throw NoDevice("plan has no device state (created with pm_plan_create_host)");
}

pm_status create(uint32_t n, uint32_t width, uint32_t limbs, const uint32_t* q, pm_plan** out, bool device) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("null output pointer");
        *out = nullptr;
        auto plan = std::make_unique<pm_plan>();
        plan->host = pm::build_plan(n, width, limbs, q, device ? pm::Tables::kDevice : pm::Tables::kHost);
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

pm_status pm_multiply_device(pm_plan* plan, const uint32_t* a, const uint32_t* b, uint32_t* c, void* stream) {
    return guarded([&] {
        require_device(plan);
        if (!a || !b || !c) throw std::invalid_argument("null operand");
        hip_check(hipSetDevice(plan->device_id), "hipSetDevice");
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : plan->stream;
        plan->run(a, b, c, PM_STAGE_OUTPUT, st);
        hip_check(hipStreamSynchronize(st), "multiply");
        plan->collect_times();
    });
}

pm_status pm_operand_create(pm_plan* plan, const uint32_t* b, pm_operand** out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("null output pointer");
        *out = nullptr;
        require_device(plan);
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
        hipStream_t st = plan->stream;
        const size_t bytes = size_t(s.n) * s.limbs * sizeof(uint32_t);
        const uint32_t ntiles = s.n_pad / 256;
        hip_check(hipMemcpyAsync(plan->io_b, b, bytes, hipMemcpyHostToDevice, st), "copy b to device");
        pm::launch_pack(s, nullptr, plan->io_b, plan->work.packed, 0, ntiles, 1, 1, st);
        pm::launch_conversion(s, plan->tables, plan->work.packed, plan->work.planes, 0, ntiles, 1, 1, st);
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
        hip_check(hipSetDevice(plan->device_id), "hipSetDevice");
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : plan->stream;
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
        if (!a || !b) throw std::invalid_argument("null operand");
        if (s.trivial && stage != PM_STAGE_OUTPUT) throw std::invalid_argument("q == 1 has no intermediate stages");
        hip_check(hipSetDevice(plan->device_id), "hipSetDevice");
        const size_t nl = size_t(s.n) * s.limbs * 4;
        hipStream_t st = plan->stream;
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
            if (stage == PM_STAGE_WORDS) {
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
