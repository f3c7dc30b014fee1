/*
 * C ABI of the MI355X (gfx950) exact negacyclic multiplier
 *     c = a * b  mod (X^N + 1, q)
 * over canonical little-endian 32-bit limb arrays (coefficient-major [N][L]).
 *
 * This is the thin layer below the drop-in C++ surface of include/fherma.h.
 * The reference exposes only that C++ surface (src/fherma.h:46-48, with
 * std::vector-backed types and exceptions); the functions here carry the same
 * three roles with plain pointers and status codes, and the fherma shim
 * (fherma-polynomial-cuda_amd/csrc/fherma_shim.cpp) turns codes into
 * std::runtime_error exactly where the reference throws.
 *
 * Threading: a plan may be used by one thread at a time.  Every entry point
 * returns only after its device work has finished.
 */
#ifndef POLYMUL_AMD_H
#define POLYMUL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pm_status {
    PM_OK = 0,
    PM_ERR_INVALID_ARGUMENT = 1,  /* bad point or input size (solve.cu:903-904, 1228) */
    PM_ERR_INSUFFICIENT_FIELDS = 2, /* no usable prime family covers 4N(q-1)^2 (solve.cu:910),
                                       e.g. q needs more than 16383 primes */
    PM_ERR_DEVICE = 3,            /* a HIP call failed (solve.cu:80-82) */
    PM_ERR_OUT_OF_MEMORY = 4,
    PM_ERR_INTERNAL = 5,
    PM_ERR_NO_DEVICE = 6          /* compute requested but no GPU / no device state */
} pm_status;

typedef struct pm_plan pm_plan;

typedef struct pm_plan_info {
    uint32_t n, width, limbs;
    uint32_t primes;        /* r: number of ~31-bit NTT fields */
    uint32_t block;         /* n1: transform split N = n1 * n2 (1 = plain negacyclic NTT) */
    uint32_t transform;     /* n2 */
    uint32_t row_width;     /* NTT row length C */
    uint32_t reference_walk;/* 1 when the primes are exactly fherma_init's walk */
    uint32_t qbits;
    uint32_t trivial;       /* 1 when q == 1 */
    uint64_t device_bytes;  /* device memory held by the plan (0 for host-only plans) */
} pm_plan_info;

/* Parameter-only initialization: replaces fherma_init (src/fherma.h:46;
 * candidates/ordered-streaming/solve.cu:900-1040).  width = bits of q,
 * limbs = ceil(width/32), q = limbs little-endian words.  Builds the prime
 * family and every table on the host, uploads them to the current device and
 * allocates all working memory.  No operand-dependent work happens here. */
pm_status pm_plan_create(uint32_t n, uint32_t width, uint32_t limbs, const uint32_t* q,
                         pm_plan** out);

/* Same plan, host tables only (no device, no HIP calls): for inspecting the
 * parameter-only plan on a machine without a GPU.  Compute entry points
 * return PM_ERR_NO_DEVICE on such a plan. */
pm_status pm_plan_create_host(uint32_t n, uint32_t width, uint32_t limbs, const uint32_t* q,
                              pm_plan** out);

/* A plan whose transform split N = n1 * n2 starts at n1 >= min_block (a power
 * of two <= min(N, 2048)); pm_plan_create is min_block = 1, and takes the
 * smallest split whose prime family suffices.  device = 0 builds host tables
 * only, like pm_plan_create_host.  No counterpart in the reference, whose
 * transform never splits: it runs each base-case kernel (n1 = 2, 4, 8 and the
 * generic n1 >= 16) at sizes where the oracle checks every output. */
pm_status pm_plan_create_split(uint32_t n, uint32_t width, uint32_t limbs, const uint32_t* q, uint32_t min_block,
                               int device, pm_plan** out);

/* Releases everything the plan holds: replaces fherma_free (src/fherma.h:48). */
void pm_plan_destroy(pm_plan* plan);

/* Host vectors in, host vector out, one product: replaces fherma_run
 * (src/fherma.h:47; candidates/ordered-streaming/solve.cu:1226-1271).  a, b, c
 * are [N][L] words; c may not alias a or b.  Includes the host->device and
 * device->host transfers, like the reference's timed window. */
pm_status pm_multiply(pm_plan* plan, const uint32_t* a, const uint32_t* b, uint32_t* c);

/* pm_multiply in two halves, for callers that prepare the output buffer while
 * the device works (the harness's fherma_run returns a newly allocated, zeroed
 * vector; the reference builds it concurrently too, output_worker.h:17-75).
 * pm_multiply_begin copies a and b in and queues every device step; it returns
 * with the device busy.  pm_multiply_finish waits and writes c (c == NULL: waits
 * and discards the result).  a and b may be released after begin returns.  In
 * between, every other product, operand or stage call on the plan fails with
 * PM_ERR_INVALID_ARGUMENT (the begun product still uses its buffers). */
pm_status pm_multiply_begin(pm_plan* plan, const uint32_t* a, const uint32_t* b);
pm_status pm_multiply_finish(pm_plan* plan, uint32_t* c);

/* Device-resident product: a, b, c are device pointers to [N][L] words on the
 * plan's device; stream is a hipStream_t (NULL = the plan's own stream).
 * No counterpart in the reference (its operands always arrive in host
 * vectors); it is the kernel-only form of fherma_run for callers whose
 * operands already live in HBM.  Synchronizes the stream before returning. */
pm_status pm_multiply_device(pm_plan* plan, const uint32_t* a, const uint32_t* b, uint32_t* c,
                             void* stream);

/* Many device-resident products, each on its own plan (its own N, q and
 * tables): plans[i] multiplies a[i] * b[i] into c[i] (device pointers to
 * [N][L] words on the plans' device).  No counterpart in the reference (it
 * multiplies one pair per call); it is the batched form for callers with many
 * products in HBM at once.  Each stage runs as one launch for every product
 * of the batch (per kernel variant), a work list mapping the launch's blocks
 * to (product, block of that product's own launch), instead of one launch per
 * stage and product: small products, which cannot fill the GPU alone and are
 * bound by launches, then share every launch.  A product whose shape has a
 * stage without a grouped kernel (q == 1, or more than 64 transform rows)
 * runs on its own, on the same stream.  Every plan must be idle and appear
 * once (its working buffers are per product); all on one device.  stream is
 * a hipStream_t (NULL = plans[0]'s own stream).  The items and work lists go
 * into a device buffer the library keeps per device (grown as needed, about
 * 0.5 KB per product, held until the process ends); batch calls on one device
 * take turns with it.  Synchronizes the stream before returning.  count == 0
 * does nothing. */
pm_status pm_multiply_device_batch(pm_plan* const* plans, const uint32_t* const* a, const uint32_t* const* b,
                                   uint32_t* const* c, size_t count, void* stream);

/* pm_multiply_device without any allocation: no counterpart in the reference.
 * It is for callers that cannot allocate device memory in the call (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The product's working buffers (packed operands, residue planes, CRT word
 * sums) come from a static workspace of PM_SCRATCH_BYTES (128 MiB) per GPU,
 * which each GPU that loads the library holds until the process ends, instead
 * of the plan's own buffers; the call makes no hipMalloc, hipFree or other
 * allocating call.  A plan whose workspace exceeds PM_SCRATCH_BYTES is
 * rejected with PM_ERR_INVALID_ARGUMENT: the reference's own point (N = 65536,
 * 1747-bit q, 125.5 MiB) fits, N = 65536 with a q of 61 limbs or more does
 * not.  Every call of this entry point on one GPU shares that GPU's workspace,
 * which keeps the last call's contents: such calls that overlap on one GPU
 * (from different threads, or queued on different streams) race and must be
 * serialized by the caller.  The plan's other entry points use the plan's own
 * buffers and do not touch the workspace. */
#define PM_SCRATCH_BYTES ((size_t)128 << 20)
pm_status pm_multiply_device_scratch(pm_plan* plan, const uint32_t* a, const uint32_t* b, uint32_t* c,
                                     void* stream);

/* Resident operand: no counterpart in the reference (its operands always
 * arrive as a pair).  For callers that multiply many a by one fixed b: b is
 * converted and transformed once and its residues stay on the device
 * (r * n_pad * 4 bytes), so each later product skips b's conversion GEMM and
 * forward transform.  An operand is tied to the plan that created it, which
 * must outlive it. */
typedef struct pm_operand pm_operand;
pm_status pm_operand_create(pm_plan* plan, const uint32_t* b, pm_operand** out);
pm_status pm_multiply_operand(pm_plan* plan, const uint32_t* a, const pm_operand* b, uint32_t* c);
pm_status pm_multiply_operand_device(pm_plan* plan, const uint32_t* a, const pm_operand* b, uint32_t* c,
                                     void* stream);
void pm_operand_destroy(pm_operand* operand);

/* Message describing the most recent failure on the calling thread. */
const char* pm_last_error(void);

pm_status pm_plan_get_info(const pm_plan* plan, pm_plan_info* info);

/* Copies min(capacity, r) primes; returns PM_OK. */
pm_status pm_plan_get_primes(const pm_plan* plan, uint32_t* primes, size_t capacity);

/* Host tables of the plan, for tests.  Identifiers: */
enum {
    PM_TABLE_PRIME_CONSTS = 1, /* [r] x {p, p^-1 mod 2^32, off_a, off_b} u32 */
    PM_TABLE_CONV_WEIGHTS = 2, /* [2][4*r_pad][conv_k] int8 */
    PM_TABLE_TWIDDLES_FWD = 3, /* [r][n2][2] u32 */
    PM_TABLE_TWIDDLES_INV = 4, /* [r][n2][2] u32 */
    PM_TABLE_CRT_WEIGHTS = 5,  /* [4*crt_words_pad][4*r_pad] int8 */
    PM_TABLE_M_MOD_Q = 6,      /* [r][L] u32: M_k = P_k mod q */
    PM_TABLE_NEG_P_MOD_Q = 7,  /* [L] u32: (-P) mod q */
    PM_TABLE_CORR_T = 8,       /* [L+2] u32 */
    PM_TABLE_CORR_E = 9,       /* [4] u32 */
    PM_TABLE_RECIPROCALS = 10, /* [r] u64: floor((2^64-1)/p_k) */
    PM_TABLE_BETA = 11,        /* [r] u32: operand-b scale */
    PM_TABLE_SHAPE = 12        /* 24 x u32: n, width, limbs, n1, n2, log2 n2, r, r_pad, conv_k,
                                  conv_rows, crt_words, crt_words_pad, crt_rows, crt_k, t_words,
                                  row_width, rows, qbits, s_shift, k_shift, mu (lo, hi), trivial,
                                  reference_walk */
};
/* Returns the byte size of a table through *bytes when out == NULL. */
pm_status pm_plan_export_table(const pm_plan* plan, int table, void* out, size_t* bytes);

/* Per-stage parity taps (GPU): runs the pipeline on host inputs a, b and
 * copies out the buffer produced by `stage`:
 *   PM_STAGE_RESIDUES  [2][r][N] u32  a mod p_k and beta_k * b mod p_k
 *   PM_STAGE_PRODUCT   [r][N]    u32  v_k = x * P_k^-1 mod p_k (x = exact
 *                                      signed negacyclic integer product)
 *   PM_STAGE_WORDS     [L+4][N]  i64  CRT word sums (signed-byte form)
 *   PM_STAGE_OUTPUT    [N][L]    u32  c
 * Returns the needed size through *bytes when out == NULL. */
enum { PM_STAGE_RESIDUES = 1, PM_STAGE_PRODUCT = 2, PM_STAGE_WORDS = 3, PM_STAGE_OUTPUT = 4 };
pm_status pm_debug_stage(pm_plan* plan, const uint32_t* a, const uint32_t* b, int stage, void* out,
                         size_t* bytes);

/* ---- Many products: the batched and served forms ----
 * No counterpart in the reference, whose fherma_run multiplies one pair at a
 * time (candidates/ordered-streaming/solve.cu:1226-1271).  A request carries its
 * own point (N, W, L, q) and host operands, like one fherma_init + fherma_run.
 * A server runs requests on `lanes` worker lanes (a compute and two copy
 * streams, pinned staging and copy threads each, reused across requests) and
 * one device arena of `device_budget_bytes`: a request's plan is built (tables
 * on the device) once its exact device footprint fits in the arena, so requests
 * are admitted by memory as earlier ones finish, and the transfers and kernels
 * of requests on different lanes overlap.  Each result is bit-identical to
 * pm_multiply on the same inputs.  A request's buffers must stay valid until it
 * completes; a request larger than the whole budget fails with
 * PM_ERR_OUT_OF_MEMORY. */
typedef struct pm_request {
    uint32_t n, width, limbs;
    const uint32_t* q;  /* [L] */
    const uint32_t* a;  /* [N][L] host */
    const uint32_t* b;  /* [N][L] host */
    uint32_t* c;        /* [N][L] host, written */
} pm_request;

typedef struct pm_request_times {
    double queue_ms;    /* submitted -> admitted (host plan built, device memory reserved) */
    double service_ms;  /* admitted -> result in c (device tables, transfers, kernels) */
    double total_ms;    /* submitted -> result in c */
} pm_request_times;

typedef struct pm_server pm_server;
pm_status pm_server_create(uint64_t device_budget_bytes, uint32_t lanes, pm_server** out);
/* Queues a request (first in, first out); returns at once with its ticket. */
pm_status pm_server_submit(pm_server* server, const pm_request* request, uint64_t* ticket);
/* Blocks until the request has finished; returns the request's own status (its
 * message through pm_last_error) and its times.  Each ticket is waited once:
 * a second wait, also one from another thread while the first still blocks,
 * fails with PM_ERR_INVALID_ARGUMENT. */
pm_status pm_server_wait(pm_server* server, uint64_t ticket, pm_request_times* times);
/* Completes every queued request, then releases the lanes and the arena. */
void pm_server_destroy(pm_server* server);

/* Every request through a temporary server, submitted largest first (by
 * N * L^2) so that small products fill the lanes at the end; statuses
 * (optional, [count]) gets each request's status; the return value is the first
 * failure (in submission order), or PM_OK. */
pm_status pm_multiply_batch(const pm_request* requests, size_t count, uint64_t device_budget_bytes, uint32_t lanes,
                            pm_status* statuses);

/* Kernel timing of the most recent device multiply (pm_multiply_device or
 * pm_multiply_operand_device), when enabled: fills up to `count` entries of ms
 * (0: operand packing, 1: conversion GEMM, 2: forward column NTT, 3: row NTT +
 * product + inverse row NTT, 4: inverse column NTT + CRT tiling, 5: CRT GEMM,
 * 6: reconstruction, 7: total), measured with HIP events on the stream the
 * kernels run on. */
pm_status pm_set_profiling(pm_plan* plan, int enabled);
pm_status pm_get_stage_times(const pm_plan* plan, float* ms, int count);

#ifdef __cplusplus
}
#endif

#endif /* POLYMUL_AMD_H */
