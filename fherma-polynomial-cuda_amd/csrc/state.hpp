// Internal state of the library (not part of the C ABI): error mapping, the
// per-worker HIP resources (Lane), the device memory arena of a server, and the
// plan and resident-operand objects behind the opaque handles of polymul.h.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "arena.hpp"
#include "chunks.hpp"
#include "copy_pool.hpp"
#include "pinned_cache.hpp"
#include "kernels.hpp"
#include "plan.hpp"
#include "polymul.h"

extern thread_local std::string g_last_error;

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct NoDevice : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char* where) {
    if (e != hipSuccess) throw DeviceError(std::string(where) + ": " + hipGetErrorString(e));
}

using pm::aligned_bytes;  // arena.hpp: device allocations in 256-byte units

// Events bracket the 7 stage launches; times are the 7 intervals plus the total.
constexpr int kEvents = 8;
constexpr int kStages = 8;

namespace pm {

// HIP resources of one sequence of host-vector products: a compute stream and two
// copy streams (all default priority), the chunk events, stage-timing events,
// pinned staging that grows to the largest product seen, and host copy workers.
// A standalone plan owns one; a server worker owns one and lends it to each plan
// it runs, so a stream of requests pays for none of this per request.
class Lane {
public:
    explicit Lane(unsigned copy_threads) {
        try {
            for (hipStream_t* st : {&stream, &copy_in, &copy_out})
                hip_check(hipStreamCreateWithFlags(st, hipStreamNonBlocking), "hipStreamCreate");
            for (int i = 0; i < kMaxChunks; ++i)
                for (hipEvent_t* e : {&in_ready[i], &out_ready[i], &reconstructed[i]})
                    hip_check(hipEventCreateWithFlags(e, hipEventDisableTiming), "hipEventCreate");
            for (hipEvent_t& e : marks) hip_check(hipEventCreate(&e), "hipEventCreate");
            pool = std::make_unique<CopyPool>(copy_threads);
        } catch (...) {
            release();
            throw;
        }
    }
    ~Lane() { release(); }
    Lane(const Lane&) = delete;
    Lane& operator=(const Lane&) = delete;

    // Copies into pinned staging peak at about 8 threads (more threads only
    // contend for the same memory bandwidth), and half the host stays free.
    static unsigned default_copy_threads() {
        const unsigned hw = std::thread::hardware_concurrency();
        return hw ? std::max(1u, std::min(8u, hw / 2)) : 4u;
    }

    void synchronize() {
        for (hipStream_t st : {stream, copy_in, copy_out})
            if (st) (void)hipStreamSynchronize(st);
    }

    // Pinned staging for products of up to `words` words per operand:
    // [2][words] in (a then b), [words] out.
    // The blocks come from (and go back to) the process-wide PinnedCache.
    void reserve(size_t words) {
        if (words <= pin_words) return;
        synchronize();
        free_pinned();
        PinnedCache& cache = PinnedCache::instance();
        pin_in = static_cast<uint32_t*>(cache.take(2 * words * sizeof(uint32_t), &pin_in_bytes));
        try {
            pin_out = static_cast<uint32_t*>(cache.take(words * sizeof(uint32_t), &pin_out_bytes));
        } catch (...) {
            free_pinned();
            throw;
        }
        pin_words = std::min(pin_in_bytes / (2 * sizeof(uint32_t)), pin_out_bytes / sizeof(uint32_t));
    }

    hipStream_t stream = nullptr, copy_in = nullptr, copy_out = nullptr;
    hipEvent_t in_ready[kMaxChunks] = {}, out_ready[kMaxChunks] = {}, reconstructed[kMaxChunks] = {};
    hipEvent_t marks[kEvents] = {};
    std::unique_ptr<CopyPool> pool;
    uint32_t* pin_in = nullptr;
    uint32_t* pin_out = nullptr;
    size_t pin_words = 0;

private:
    size_t pin_in_bytes = 0, pin_out_bytes = 0;

    void free_pinned() {
        PinnedCache& cache = PinnedCache::instance();
        if (pin_in) cache.give(pin_in, pin_in_bytes);
        if (pin_out) cache.give(pin_out, pin_out_bytes);
        pin_in = pin_out = nullptr;
        pin_in_bytes = pin_out_bytes = pin_words = 0;
    }
    void release() {
        synchronize();
        free_pinned();
        for (int i = 0; i < kMaxChunks; ++i)
            for (hipEvent_t e : {in_ready[i], out_ready[i], reconstructed[i]})
                if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : marks)
            if (e) (void)hipEventDestroy(e);
        for (hipStream_t st : {stream, copy_in, copy_out})
            if (st) (void)hipStreamDestroy(st);
        stream = copy_in = copy_out = nullptr;
    }
};

}  // namespace pm

struct pm_plan {
    pm::HostPlan host;
    bool device = false;
    int device_id = 0;
    pm::Lane* lane = nullptr;              // streams, events, pinned staging, copy workers
    std::unique_ptr<pm::Lane> own_lane;    // set for a standalone plan
    // Device memory: a standalone plan hipMallocs each buffer; a plan run by a
    // server carves them, in order, out of one block of the server's arena.
    char* block = nullptr;
    size_t block_size = 0, block_used = 0;
    std::vector<void*> allocations;
    pm::DeviceTables tables;
    pm::DeviceWork work;
    uint32_t* io_a = nullptr;  // device copies for the host-vector entry point
    uint32_t* io_b = nullptr;
    uint32_t* io_c = nullptr;
    uint64_t device_bytes = 0;
    uint32_t* dev_beta = nullptr;  // device-built beta (Tables::kDevice), for table export
    bool pending = false;          // multiply_begin done, multiply_finish not yet
    bool profiling = false;
    float stage_ms[kStages] = {};

    template <class T>
    T* alloc(size_t count) {
        const size_t bytes = aligned_bytes(count * sizeof(T));
        void* p = nullptr;
        if (block) {
            if (block_used + bytes > block_size) throw std::logic_error("plan: device block smaller than its footprint");
            p = block + block_used;
            block_used += bytes;
        } else {
            hipError_t e = hipMalloc(&p, bytes);
            if (e != hipSuccess) {
                if (e == hipErrorOutOfMemory) throw std::bad_alloc();
                hip_check(e, "hipMalloc");
            }
            allocations.push_back(p);
        }
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
            (void)hipSetDevice(device_id);
            if (lane) lane->synchronize();
            for (void* p : allocations) (void)hipFree(p);
        }
    }

    pm::Chunking chunking() const {
        const pm::Shape& s = host.shape;
        return pm::make_chunking(s.n_pad, s.n, s.limbs, s.row_width, s.rows);
    }
    // First half of a host-vector product, per coefficient chunk (csrc/chunks.hpp):
    // the copy workers gather the chunk's rows of a and b (one job) from the
    // caller's pageable vectors into pinned staging as [a rows | b rows], one H2D
    // carries both on a copy stream, and the compute stream waits on an event
    // before packing and converting that chunk -- the host copy of chunk k+1
    // overlaps the H2D of chunk k, the conversion of chunk k-1 and this
    // thread's HIP calls that queue chunk k (posted to the workers first, so
    // the copy does not wait for those calls).  When the
    // chunks are column strips, a strip's forward column stages follow the
    // chunk that completes it, under the later chunks' transfers; the row stage
    // needs every coefficient.  On the way out, per chunk: the inverse column
    // stages of the strips it starts, its CRT GEMM and its reconstruction, which
    // reaches pinned memory while the next chunk is computed -- written there by
    // the kernel itself when it stores each word once (short q), else by a D2H
    // on the second copy stream.  Each DMA and each hand-off between the compute
    // stream and a copy stream costs some 10-40 us on this link, which is why
    // both are kept few.  Staging, the device copies of the operands and of c
    // hold the chunks one after another (chunk-major rows).
    // Returns with the device busy; multiply_finish collects c.
    // b == nullptr: the b residues come from `bplanes_in` (a resident operand).
    void multiply_begin(const uint32_t* a, const uint32_t* b, const uint32_t* bplanes_in) {
        const pm::Shape& s = host.shape;
        const uint32_t nops = b ? 2 : 1;
        const uint32_t* bplanes = b ? b_planes() : bplanes_in;
        const size_t L = s.limbs, row_bytes = L * sizeof(uint32_t);
        if (pending) throw std::invalid_argument("the plan's previous product has not been finished");
        if (s.trivial) {
            pending = true;
            return;
        }
        lane->reserve(size_t(s.n) * L);
        hipStream_t stream = lane->stream, copy_in = lane->copy_in, copy_out = lane->copy_out;
        pm::CopyPool* pool = lane->pool.get();
        pool->set_active(true);  // workers spin between the chunk jobs of this product
        struct Idle {
            pm::CopyPool* pool;
            bool armed = true;
            ~Idle() {
                if (armed) pool->set_active(false);
            }
        } idle{pool};
        // A failure while queuing leaves steps already queued still running on
        // buffers the next call would reuse: wait for them before rethrowing.
        try {
            const pm::Chunking chunks = chunking();
            // Copy jobs: a job is one chunk's rows [r0, r1) of operands
            // [op0, op1).  Nothing overlaps the first job's host copy, so the
            // first chunk goes in pieces: with one run (tile ranges) its a rows
            // in two halves, then (two operands) its b rows, each with its own
            // H2D, which then runs while the next piece is copied; with
            // several runs (strips) a's rows, then b's rows.
            struct Job {
                uint32_t ch, op0, op1;
                size_t r0, r1;
            };
            Job jobs[pm::kMaxChunks + 2];
            uint32_t njobs = 0;
            for (uint32_t ch = 0; ch < chunks.count; ++ch) {
                const size_t rows = chunks.rows_of(ch);
                // (halves only when a's rows are 2 MB or more: below that the
                // extra transfer's hand-off costs more than the copy it hides)
                if (ch == 0 && chunks.runs(0) == 1 && rows * row_bytes >= (size_t(2) << 20)) {
                    jobs[njobs++] = {0, 0, 1, 0, rows / 2};
                    jobs[njobs++] = {0, 0, 1, rows / 2, rows};
                    if (nops == 2) jobs[njobs++] = {0, 1, 2, 0, rows};
                } else if (ch == 0 && nops == 2) {
                    jobs[njobs++] = {0, 0, 1, 0, rows};
                    jobs[njobs++] = {0, 1, 2, 0, rows};
                } else {
                    jobs[njobs++] = {ch, 0, nops, 0, rows};
                }
            }
            // A job's rows, at word offset nops*local*L in both staging and
            // io_a (the chunk's a rows, then its b rows).  A job of part of
            // the rows (one run, one operand) is one segment.
            pm::CopyPool::Segment segs[2 * pm::kMaxRuns];
            auto post = [&](const Job& jb) {
                const size_t rows = chunks.rows_of(jb.ch), base = size_t(nops) * chunks.local[jb.ch] * L;
                const uint32_t runs = chunks.runs(jb.ch);
                if (runs > pm::kMaxRuns) throw std::logic_error("a chunk of more runs than kMaxRuns");
                size_t k = 0;
                for (uint32_t op = jb.op0; op < jb.op1; ++op)
                    for (uint32_t g = 0; g < runs; ++g) {
                        size_t global0, local0, len;
                        chunks.run(jb.ch, g, s.n, global0, local0, len);
                        if (jb.r1 - jb.r0 < rows) {  // one run: its rows [r0, r1)
                            global0 += jb.r0, local0 += jb.r0, len = jb.r1 - jb.r0;
                        }
                        segs[k++] = {lane->pin_in + base + (op * rows + local0 - chunks.local[jb.ch]) * L,
                                     (op ? b : a) + global0 * L, len * row_bytes};
                    }
                pool->post(segs, k);
            };
            // Two copy streams remove the hand-off between consecutive chunks'
            // transfers, but run them two at a time, so each chunk, and its
            // conversion, lands later.  That pays while a chunk's conversion is
            // short against its transfer: conversion at about 2.6e15 INT8 ops/s
            // of 4 * conv_rows * conv_k ops per coefficient, transfer at about
            // 57 GB/s of 8 * L bytes; under half means conv_rows * conv_k <
            // 45600 * L.  (N = 32768 with 868 limbs, ratio 0.33: 7.06 against
            // 7.27 ms per product with one stream; N = 65536 with 1747 limbs,
            // ratio 0.64: 28.2 against 27.5 ms.)
            const bool two_streams = uint64_t(s.conv_rows) * s.conv_k < uint64_t(45600) * L;
            post(jobs[0]);
            for (uint32_t j = 0; j < njobs; ++j) {
                const Job jb = jobs[j];
                const uint32_t ch = jb.ch;
                const size_t rows = chunks.rows_of(ch), base = size_t(nops) * chunks.local[ch] * L;
                pool->wait();
                const size_t at = base + (jb.op0 * rows + jb.r0) * L;
                const size_t bytes = jb.r1 - jb.r0 < rows ? (jb.r1 - jb.r0) * row_bytes
                                                          : (jb.op1 - jb.op0) * rows * row_bytes;
                // With two_streams, odd chunks go on the output copy stream, whose
                // transfers all come after the last input, so that consecutive
                // chunks' transfers do not queue behind each other's completion.
                // (A fourth stream per lane would put every lane's compute stream
                // on one of the runtime's round-robin hardware queues, serializing
                // concurrent plans.)
                hipStream_t in = (two_streams && (ch & 1)) ? copy_out : copy_in;
                hip_check(hipMemcpyAsync(io_a + at, lane->pin_in + at, bytes, hipMemcpyHostToDevice, in),
                          "H2D inputs");
                // the workers copy the next job while this thread queues the
                // chunk's kernels
                if (j + 1 < njobs) post(jobs[j + 1]);
                if (jb.op1 < nops || jb.r1 < rows) continue;  // more of the chunk is still to come
                hip_check(hipEventRecord(lane->in_ready[ch], in), "hipEventRecord");
                hip_check(hipStreamWaitEvent(stream, lane->in_ready[ch], 0), "hipStreamWaitEvent");
                const uint32_t* xa = io_a + base;  // pack reads the chunk's k-th tile at rows k*256..
                const uint32_t* xb = xa + rows * L;
                pm::launch_pack(s, xa, xb, work.packed, chunks.set[ch], 0, nops, stream);
                pm::launch_conversion(s, tables, work.packed, work.planes, chunks.set[ch], 0, nops, stream);
                if (chunks.strips && chunks.fwd_end[ch] > chunks.fwd_begin[ch])
                    pm::launch_ntt_forward_columns(s, tables, work.planes, 0, nops, stream, chunks.fwd_begin[ch],
                                                   chunks.fwd_end[ch]);
            }
            if (!chunks.strips) pm::launch_ntt_forward_columns(s, tables, work.planes, 0, nops, stream);
            pm::launch_ntt_rows_product(s, tables, work.planes, bplanes, stream);
            if (!chunks.strips) pm::launch_ntt_inverse_columns(s, tables, work.planes, work.rtiles, stream);
            // Output half, per chunk: the inverse column stages of the strips it
            // starts, its CRT GEMM, its reconstruction, then its D2H on the copy
            // stream, which overlaps the next chunk's GEMM.
            const bool direct = pm::reconstruct_writes_once(s);
            for (uint32_t ch = 0; ch < chunks.count; ++ch) {
                const size_t rows = chunks.rows_of(ch), at = size_t(chunks.local[ch]) * L;
                if (chunks.strips && chunks.inv_end[ch] > chunks.inv_begin[ch])
                    pm::launch_ntt_inverse_columns(s, tables, work.planes, work.rtiles, stream, chunks.inv_begin[ch],
                                                   chunks.inv_end[ch]);
                pm::launch_crt(s, tables, work.rtiles, work.words, chunks.set[ch], stream);
                pm::launch_reconstruct(s, tables, work.words, work.t_words, (direct ? lane->pin_out : io_c) + at,
                                       chunks.set[ch], stream);
                if (direct) {
                    hip_check(hipEventRecord(lane->out_ready[ch], stream), "hipEventRecord");
                    continue;
                }
                // With two_streams the D2Hs alternate too (the input stream is idle
                // by now), so that a chunk's D2H does not wait behind the previous
                // one's completion event.  Where a chunk's conversion is long
                // against its transfer the chunks are large, and then the host's
                // copy of each into the caller's vector, whose fresh pages fault on
                // first write, is what sets the pace: D2Hs landing in pairs only
                // build a backlog there, so they stay on one stream.
                hipStream_t out = (two_streams && (ch & 1)) ? copy_in : copy_out;
                hip_check(hipEventRecord(lane->reconstructed[ch], stream), "hipEventRecord");
                hip_check(hipStreamWaitEvent(out, lane->reconstructed[ch], 0), "hipStreamWaitEvent");
                hip_check(hipMemcpyAsync(lane->pin_out + at, io_c + at, rows * row_bytes, hipMemcpyDeviceToHost, out),
                          "D2H c");
                hip_check(hipEventRecord(lane->out_ready[ch], out), "hipEventRecord");
            }
        } catch (...) {
            pool->wait();  // a posted copy still writes into the staging
            lane->synchronize();
            throw;
        }
        pending = true;  // only once every step is queued: a failure above leaves the plan usable
        idle.armed = false;  // the pool keeps spinning until multiply_finish
    }

    // Second half: waits for each chunk's results and scatters them into c
    // (c == nullptr: waits and discards the result).
    void multiply_finish(uint32_t* c) {
        if (!pending) throw std::invalid_argument("no product has been begun on this plan");
        pending = false;
        const pm::Shape& s = host.shape;
        const size_t L = s.limbs, row_bytes = L * sizeof(uint32_t);
        if (s.trivial) {
            if (c) std::memset(c, 0, size_t(s.n) * row_bytes);
            return;
        }
        struct Idle {
            pm::CopyPool* pool;
            ~Idle() { pool->set_active(false); }
        } idle{lane->pool.get()};
        const pm::Chunking chunks = chunking();
        for (uint32_t ch = 0; ch < chunks.count; ++ch) {
            hip_check(hipEventSynchronize(lane->out_ready[ch]), "D2H c");
            if (!c) continue;
            const uint32_t runs = std::min(chunks.runs(ch), pm::kMaxRuns);  // multiply_begin checked it
            pm::CopyPool::Segment segs[pm::kMaxRuns];
            for (uint32_t g = 0; g < runs; ++g) {
                size_t global0, local0, len;
                chunks.run(ch, g, s.n, global0, local0, len);
                segs[g] = {c + global0 * L, lane->pin_out + local0 * L, len * row_bytes};
            }
            lane->pool->copy(segs, runs);
        }
        // Nothing follows the last chunk's event on any stream, so the streams
        // are idle here (a stream synchronize would cost 20-40 us for nothing).
    }

    void multiply_host(const uint32_t* a, const uint32_t* b, const uint32_t* bplanes_in, uint32_t* c) {
        multiply_begin(a, b, bplanes_in);
        multiply_finish(c);
    }

    void make_device() {
        hip_check(hipGetDevice(&device_id), "hipGetDevice");
        device = true;
        if (!lane) {
            own_lane = std::make_unique<pm::Lane>(pm::Lane::default_copy_threads());
            lane = own_lane.get();
        }
        const pm::Shape& s = host.shape;
        const size_t nl = size_t(s.n) * s.limbs;
        io_a = alloc<uint32_t>(2 * nl);  // a then b; the host path stages chunks as [a rows | b rows]
        io_b = io_a + nl;
        io_c = alloc<uint32_t>(nl);
        if (s.trivial) return;
        pm::init_scratch();  // per-device address of pm_multiply_device_scratch's static workspace
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
        const size_t half = size_t(std::max(s.r_pad, pm::scratch_rows(s))) * s.n_pad;
        work.planes = alloc<uint32_t>(2 * half);
        work.rtiles = reinterpret_cast<int8_t*>(work.planes + half);
        work.t_words = work.planes;
        work.words = alloc<int64_t>(pm::word_sum_elements(s));
        // pinned staging of the host entry points is sized here, with the rest
        // of the plan, not inside the first product
        lane->reserve(nl);
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
        pm::launch_build_tables(s, tb, lane->stream);
        hip_check(hipStreamSynchronize(lane->stream), "build tables");
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
    uint32_t* b_planes() const { return b_planes(work); }
    uint32_t* b_planes(const pm::DeviceWork& w) const { return w.planes + size_t(host.shape.r_pad) * host.shape.n_pad; }

    // Runs stages up to and including `last` (PM_STAGE_*) on device operands;
    // b == nullptr takes the b residues from `bplanes_in` (a resident operand).
    void run(const uint32_t* a, const uint32_t* b, uint32_t* c, int last, hipStream_t st,
             const uint32_t* bplanes_in = nullptr) {
        run_with(work, a, b, c, last, st, bplanes_in);
    }

    // The same stages on the working buffers `w` (the plan's own, or the
    // static workspace of pm_multiply_device_scratch).
    void run_with(const pm::DeviceWork& w, const uint32_t* a, const uint32_t* b, uint32_t* c, int last,
                  hipStream_t st, const uint32_t* bplanes_in = nullptr) {
        const pm::Shape& s = host.shape;
        if (s.trivial) {
            pm::launch_zero(c, size_t(s.n) * s.limbs, st);
            return;
        }
        auto mark = [&](int i) {
            if (profiling) hip_check(hipEventRecord(lane->marks[i], st), "hipEventRecord");
        };
        const pm::TileSet all = pm::all_tiles(s);
        const uint32_t nops = b ? 2 : 1;
        const uint32_t* bplanes = b ? b_planes(w) : bplanes_in;
        mark(0);
        pm::launch_pack(s, a, b, w.packed, all, 0, nops, st);
        mark(1);
        pm::launch_conversion(s, tables, w.packed, w.planes, all, 0, nops, st);
        mark(2);
        if (last == PM_STAGE_RESIDUES) return;
        pm::launch_ntt_forward_columns(s, tables, w.planes, 0, nops, st);
        mark(3);
        pm::launch_ntt_rows_product(s, tables, w.planes, bplanes, st);
        mark(4);
        pm::launch_ntt_inverse_columns(s, tables, w.planes, w.rtiles, st);
        mark(5);
        if (last == PM_STAGE_PRODUCT) return;
        pm::launch_crt(s, tables, w.rtiles, w.words, all, st);
        mark(6);
        if (last == PM_STAGE_WORDS) return;
        pm::launch_reconstruct(s, tables, w.words, w.t_words, c, all, st);
        mark(7);
    }

    void collect_times() {
        if (!profiling || host.shape.trivial) return;
        for (int i = 0; i < kEvents - 1; ++i)
            hip_check(hipEventElapsedTime(&stage_ms[i], lane->marks[i], lane->marks[i + 1]), "hipEventElapsedTime");
        hip_check(hipEventElapsedTime(&stage_ms[kEvents - 1], lane->marks[0], lane->marks[kEvents - 1]),
                  "hipEventElapsedTime");
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
            (void)hipSetDevice(plan->device_id);
            (void)hipFree(planes);
        }
    }
};


// Device bytes pm_plan::make_device allocates for this host plan (the same
// sequence of 256-byte-rounded buffers); a server reserves exactly this much.
inline size_t plan_footprint(const pm::HostPlan& h) {
    const pm::Shape& s = h.shape;
    const size_t nl = size_t(s.n) * s.limbs * sizeof(uint32_t);
    size_t t = aligned_bytes(2 * nl) + aligned_bytes(nl);  // io_a with io_b, io_c
    if (s.trivial) return t;
    const size_t r = s.r;
    t += aligned_bytes(r * sizeof(pm::PrimeConst)) + aligned_bytes(2 * r * 4) + aligned_bytes(r * 8) +
         aligned_bytes(r * s.limbs * 4) + aligned_bytes(r * 4);
    t += aligned_bytes(size_t(2) * s.conv_rows * s.conv_k) + 2 * aligned_bytes(r * s.n2 * 2 * 4) +
         aligned_bytes(size_t(s.crt_rows) * s.crt_k);
    t += 2 * aligned_bytes(size_t(s.limbs) * 4) + aligned_bytes(size_t(s.t_words) * 4) + aligned_bytes(4 * 4);
    const size_t half = size_t(std::max(s.r_pad, pm::scratch_rows(s))) * s.n_pad;
    t += aligned_bytes(pm::packed_operand_bytes(s)) + aligned_bytes(2 * half * 4) +
         aligned_bytes(pm::word_sum_elements(s) * 8);
    return t;
}

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
    } catch (const pm::InsufficientFields& e) {
        g_last_error = e.what();
        return PM_ERR_INSUFFICIENT_FIELDS;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return PM_ERR_INTERNAL;
    }
}

inline void require_device(const pm_plan* plan) {
    if (!plan) throw std::invalid_argument("null plan");
    if (!plan->device) throw NoDevice("plan has no device state (created with pm_plan_create_host)");
}

// Entry points that use the plan's device buffers wait for a product begun
// with pm_multiply_begin to be finished first: its kernels still use them.
inline void require_idle(const pm_plan* plan) {
    if (plan->pending) throw std::invalid_argument("the plan's begun product has not been finished");
}

