// gfx950 kernels of the exact negacyclic multiplier c = a*b mod (X^N+1, q).
//
// Pipeline (per product; every stage is integer-only):
//   1. conversion_gemm      limb bytes x signed byte weights on INT8 MFMA
//                           (v_mfma_i32_16x16x64_i8), REDC -> residues mod p_k
//                           [reference: tensor_product<true>, solve.cu:143-250]
//   2. ntt_columns_forward  merged-twist negacyclic NTT stages whose butterfly
//                           distance spans whole rows  [high_stages, solve.cu:280-382]
//   3. ntt_rows_product     remaining forward stages, base-case product mod
//                           (X^n1 - zeta), inverse stages, all inside one row
//                           [middle_stages*, solve.cu:385-570]
//   4. ntt_columns_inverse  inverse row-spanning stages -> v_k = x P_k^-1 mod p_k
//                           [high_stages<true>, solve.cu:317-382]
//   5. crt_gemm             residue bytes x signed digits of M_k and of
//                           floor((2^64-1)/p_k) on INT8 MFMA -> word sums of
//                           T = sum v_k M_k and of the quotient sum E
//                           [tensor_product<false>, solve.cu:143-250; CFS:1872-1881]
//   6. reconstruct          K = round(E/2^64), T += K*((-P) mod q), carries,
//                           Barrett reduction by an arbitrary q
//                           [reconstruct_direct, solve.cu:667-716, generalized]
//
// Signed INT8 MFMA: data bytes u in [0,255] enter as u-128 (XOR 0x80); the
// weights are signed base-256 digits; the constant 128*sum(weights) is folded
// into per-prime offsets (conversion) and a fixed big-integer correction (CRT).
#include "kernels.hpp"

#include <stdexcept>
#include <string>

namespace pm {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr uint32_t kBM = 128;     // GEMM rows per block (weight rows)
constexpr uint32_t kBN = 128;     // GEMM columns per block (coefficients)
constexpr uint32_t kBK = 64;      // bytes of K per MFMA step
constexpr uint32_t kGemmThreads = 256;

void check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

__device__ __forceinline__ uint32_t add_mod(uint32_t a, uint32_t b, uint32_t p) {
    uint32_t t = a + b;
    return t >= p ? t - p : t;
}

__device__ __forceinline__ uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t p) {
    return a >= b ? a - b : a + p - b;
}

// Shoup: wq = floor(w*2^32/p); for a < 2^32, a*w - hi(a*wq)*p lies in [0, 2p).
__device__ __forceinline__ uint32_t mul_shoup(uint32_t a, uint32_t w, uint32_t wq, uint32_t p) {
    uint32_t q = __umulhi(a, wq);
    uint32_t r = a * w - q * p;
    return r >= p ? r - p : r;
}

// Subtractive Montgomery REDC of t < p*2^32: (t - m p)/2^32, m = t p^-1 mod 2^32.
__device__ __forceinline__ uint32_t redc(uint64_t t, uint32_t p, uint32_t pinv) {
    uint32_t m = uint32_t(t) * pinv;
    uint32_t mh = __umulhi(m, p);
    uint32_t hi = uint32_t(t >> 32);
    uint32_t r = hi - mh;
    return hi < mh ? r + p : r;
}

__device__ __forceinline__ uint32_t mont(uint32_t a, uint32_t b, uint32_t p, uint32_t pinv) {
    return redc(uint64_t(a) * b, p, pinv);
}

// REDC of a signed t with |t| < p*2^32; result t*2^-32 mod p in [0, p).
__device__ __forceinline__ uint32_t redc_signed(int64_t t, uint32_t p, uint32_t pinv) {
    uint32_t m = uint32_t(uint64_t(t)) * pinv;
    int64_t r = (t >> 32) - int64_t(__umulhi(m, p)) + int64_t(p);
    if (r < 0) r += p;
    if (r >= int64_t(p)) r -= p;
    return uint32_t(r);
}

// LDS tile: rows of 64 bytes, 16-byte chunks XOR-swizzled by row/4 so the 16
// rows one MFMA fragment reads land in distinct bank groups.
__device__ __forceinline__ uint32_t lds_offset(uint32_t row, uint32_t chunk) {
    return row * kBK + (((chunk ^ (row >> 2)) & 3u) << 4);
}

__device__ __forceinline__ int64_t combine4(const v4i& v) {
    return int64_t(v[0]) + int64_t(v[1]) * 256 + int64_t(v[2]) * 65536 + int64_t(v[3]) * 16777216;
}

// One 64-byte K step of the 128x128 block tile: wave (wm, wn) owns a 64x64
// quadrant = 4x4 MFMA 16x16 tiles.  A rows = weight rows, B rows = coefficients.
__device__ __forceinline__ void mfma_step(const int8_t* lds_a, const int8_t* lds_b, uint32_t lane,
                                          uint32_t wm, uint32_t wn, v4i (&acc)[4][4]) {
    v4i fa[4], fb[4];
#pragma unroll
    for (uint32_t mi = 0; mi < 4; ++mi) {
        uint32_t row = wm * 64 + mi * 16 + (lane & 15);
        fa[mi] = *reinterpret_cast<const v4i*>(lds_a + lds_offset(row, lane >> 4));
    }
#pragma unroll
    for (uint32_t ni = 0; ni < 4; ++ni) {
        uint32_t col = wn * 64 + ni * 16 + (lane & 15);
        fb[ni] = *reinterpret_cast<const v4i*>(lds_b + lds_offset(col, lane >> 4));
    }
#pragma unroll
    for (uint32_t mi = 0; mi < 4; ++mi)
#pragma unroll
        for (uint32_t ni = 0; ni < 4; ++ni)
            acc[mi][ni] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[mi], fb[ni], acc[mi][ni], 0, 0, 0);
}

// Weight tile 128 rows x 64 bytes: 2 x 16 bytes per thread.
__device__ __forceinline__ void load_weights(const int8_t* A, uint32_t lda, uint32_t k0, uint32_t tid,
                                             uint4 (&ra)[2]) {
#pragma unroll
    for (uint32_t s = 0; s < 2; ++s) {
        uint32_t idx = tid + kGemmThreads * s, row = idx >> 2, chunk = idx & 3;
        ra[s] = *reinterpret_cast<const uint4*>(A + size_t(row) * lda + k0 + chunk * 16);
    }
}

__device__ __forceinline__ void store_weights(int8_t* lds_a, uint32_t tid, const uint4 (&ra)[2]) {
#pragma unroll
    for (uint32_t s = 0; s < 2; ++s) {
        uint32_t idx = tid + kGemmThreads * s, row = idx >> 2, chunk = idx & 3;
        *reinterpret_cast<uint4*>(lds_a + lds_offset(row, chunk)) = ra[s];
    }
}

// ---------------------------------------------------------------------------
// 1. conversion GEMM: G[4k+d][i] = sum_j digit_d(lambda*256^j mod p_k) * (byte_j(x_i)-128)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void conversion_gemm(const int8_t* __restrict__ weights,
                                                       const uint32_t* __restrict__ a,
                                                       const uint32_t* __restrict__ b,
                                                       const PrimeConst* __restrict__ consts,
                                                       uint32_t* __restrict__ planes, uint32_t n,
                                                       uint32_t limbs, uint32_t r, uint32_t conv_rows,
                                                       uint32_t conv_k) {
    __shared__ __attribute__((aligned(16))) int8_t lds_a[kBM * kBK];
    __shared__ __attribute__((aligned(16))) int8_t lds_b[kBN * kBK];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t wm = wave >> 1, wn = wave & 1;
    const uint32_t n0 = blockIdx.x * kBN, m0 = blockIdx.y * kBM, op = blockIdx.z;
    const int8_t* A = weights + (size_t(op) * conv_rows + m0) * conv_k;
    const uint32_t* src = op ? b : a;
    const uint32_t kbytes = 4 * limbs;

    uint4 ra[2];
    uint32_t rb[8];
    auto load = [&](uint32_t k0) {
        load_weights(A, conv_k, k0, tid, ra);
#pragma unroll
        for (uint32_t s = 0; s < 8; ++s) {
            uint32_t idx = tid + kGemmThreads * s, coef = idx >> 4, dw = idx & 15;
            uint32_t kb = k0 + 4 * dw, gc = n0 + coef;
            uint32_t v = (gc < n && kb < kbytes) ? src[size_t(gc) * limbs + (kb >> 2)] : 0u;
            rb[s] = v ^ 0x80808080u;
        }
    };
    auto store = [&]() {
        store_weights(lds_a, tid, ra);
#pragma unroll
        for (uint32_t s = 0; s < 8; ++s) {
            uint32_t idx = tid + kGemmThreads * s, coef = idx >> 4, dw = idx & 15;
            *reinterpret_cast<uint32_t*>(lds_b + lds_offset(coef, dw >> 2) + (dw & 3) * 4) = rb[s];
        }
    };

    v4i acc[4][4];
#pragma unroll
    for (uint32_t mi = 0; mi < 4; ++mi)
#pragma unroll
        for (uint32_t ni = 0; ni < 4; ++ni) acc[mi][ni] = v4i{0, 0, 0, 0};

    load(0);
    for (uint32_t k0 = 0; k0 < conv_k; k0 += kBK) {
        __syncthreads();
        store();
        __syncthreads();
        if (k0 + kBK < conv_k) load(k0 + kBK);
        mfma_step(lds_a, lds_b, lane, wm, wn, acc);
    }

    // Lane holds rows 4*(lane/16)+j (j = digit) of each 16-row tile: one prime.
#pragma unroll
    for (uint32_t mi = 0; mi < 4; ++mi) {
        const uint32_t kp = (m0 + wm * 64 + mi * 16) / 4 + (lane >> 4);
        if (kp >= r) continue;
        const PrimeConst pc = consts[kp];
        const int64_t off = op ? pc.off_b : pc.off_a;
        uint32_t* out = planes + (size_t(op) * r + kp) * n;
#pragma unroll
        for (uint32_t ni = 0; ni < 4; ++ni) {
            const uint32_t coef = n0 + wn * 64 + ni * 16 + (lane & 15);
            if (coef >= n) continue;
            out[coef] = redc_signed(combine4(acc[mi][ni]) + off, pc.p, pc.pinv);
        }
    }
}

// ---------------------------------------------------------------------------
// 5. CRT GEMM: G[c][i] = sum_{k,d} w[c][4k+d] * (byte_d(v_{k,i}) - 128),
//    words[j][i] = sum_t 256^t G[4j+t][i]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void crt_gemm(const int8_t* __restrict__ weights,
                                                const uint32_t* __restrict__ plane,
                                                int64_t* __restrict__ words, uint32_t n, uint32_t r,
                                                uint32_t crt_k, uint32_t crt_words) {
    __shared__ __attribute__((aligned(16))) int8_t lds_a[kBM * kBK];
    __shared__ __attribute__((aligned(16))) int8_t lds_b[kBN * kBK];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t wm = wave >> 1, wn = wave & 1;
    const uint32_t n0 = blockIdx.x * kBN, m0 = blockIdx.y * kBM;
    const int8_t* A = weights + size_t(m0) * crt_k;

    uint4 ra[2];
    uint32_t rb[8];
    auto load = [&](uint32_t k0) {
        load_weights(A, crt_k, k0, tid, ra);
#pragma unroll
        for (uint32_t s = 0; s < 8; ++s) {
            uint32_t idx = tid + kGemmThreads * s, kk = idx >> 7, coef = idx & 127;
            uint32_t k = k0 / 4 + kk, gc = n0 + coef;
            uint32_t v = (k < r && gc < n) ? plane[size_t(k) * n + gc] : 0u;
            rb[s] = v ^ 0x80808080u;
        }
    };
    auto store = [&]() {
        store_weights(lds_a, tid, ra);
#pragma unroll
        for (uint32_t s = 0; s < 8; ++s) {
            uint32_t idx = tid + kGemmThreads * s, kk = idx >> 7, coef = idx & 127;
            *reinterpret_cast<uint32_t*>(lds_b + lds_offset(coef, kk >> 2) + (kk & 3) * 4) = rb[s];
        }
    };

    v4i acc[4][4];
#pragma unroll
    for (uint32_t mi = 0; mi < 4; ++mi)
#pragma unroll
        for (uint32_t ni = 0; ni < 4; ++ni) acc[mi][ni] = v4i{0, 0, 0, 0};

    load(0);
    for (uint32_t k0 = 0; k0 < crt_k; k0 += kBK) {
        __syncthreads();
        store();
        __syncthreads();
        if (k0 + kBK < crt_k) load(k0 + kBK);
        mfma_step(lds_a, lds_b, lane, wm, wn, acc);
    }

#pragma unroll
    for (uint32_t mi = 0; mi < 4; ++mi) {
        const uint32_t word = (m0 + wm * 64 + mi * 16) / 4 + (lane >> 4);
        if (word >= crt_words) continue;
#pragma unroll
        for (uint32_t ni = 0; ni < 4; ++ni) {
            const uint32_t coef = n0 + wn * 64 + ni * 16 + (lane & 15);
            if (coef >= n) continue;
            words[size_t(word) * n + coef] = combine4(acc[mi][ni]);
        }
    }
}

// ---------------------------------------------------------------------------
// 2-4. NTT.  Flat index i = row*C + col.  Forward: Cooley-Tukey with the
// negacyclic twist merged into psi^bitrev twiddles, m = 1, 2, .., n2/2 groups
// at flat distance N/(2m) down to n1.  Inverse: Gentleman-Sande, mirror order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void ntt_columns_forward(uint32_t* __restrict__ planes,
                                                           const PrimeConst* __restrict__ consts,
                                                           const uint2* __restrict__ tw, uint32_t n,
                                                           uint32_t n2, uint32_t C, uint32_t rows,
                                                           uint32_t cw, uint32_t r) {
    extern __shared__ uint32_t tile[];  // [rows][cw]
    const uint32_t col0 = blockIdx.x * cw, k = blockIdx.y, op = blockIdx.z;
    const uint32_t p = consts[k].p;
    uint32_t* data = planes + (size_t(op) * r + k) * n;
    const uint2* w = tw + size_t(k) * n2;
    const uint32_t total = rows * cw;
    for (uint32_t t = threadIdx.x; t < total; t += blockDim.x)
        tile[t] = data[size_t(t / cw) * C + col0 + t % cw];
    __syncthreads();
    for (uint32_t m = 1, half = rows / 2; m < rows; m *= 2, half /= 2) {
        for (uint32_t t = threadIdx.x; t < total / 2; t += blockDim.x) {
            uint32_t pr = t / cw, c = t % cw;
            uint32_t g = pr / half, jr = pr % half, row = 2 * half * g + jr;
            uint2 tw_g = w[m + g];
            uint32_t u = tile[row * cw + c];
            uint32_t v = mul_shoup(tile[(row + half) * cw + c], tw_g.x, tw_g.y, p);
            tile[row * cw + c] = add_mod(u, v, p);
            tile[(row + half) * cw + c] = sub_mod(u, v, p);
        }
        __syncthreads();
    }
    for (uint32_t t = threadIdx.x; t < total; t += blockDim.x)
        data[size_t(t / cw) * C + col0 + t % cw] = tile[t];
}

__global__ __launch_bounds__(256) void ntt_rows_product(uint32_t* __restrict__ planes,
                                                        const PrimeConst* __restrict__ consts,
                                                        const uint2* __restrict__ twf,
                                                        const uint2* __restrict__ twi, uint32_t n,
                                                        uint32_t n1, uint32_t n2, uint32_t C,
                                                        uint32_t rows, uint32_t r) {
    extern __shared__ uint32_t lds[];  // A[C], B[C]
    uint32_t* A = lds;
    uint32_t* B = lds + C;
    const uint32_t row = blockIdx.x, k = blockIdx.y;
    const uint32_t p = consts[k].p, pinv = consts[k].pinv;
    uint32_t* pa = planes + size_t(k) * n + size_t(row) * C;
    const uint32_t* pb = planes + (size_t(r) + k) * n + size_t(row) * C;
    const uint2* wf = twf + size_t(k) * n2;
    const uint2* wi = twi + size_t(k) * n2;
    for (uint32_t t = threadIdx.x; t < C; t += blockDim.x) {
        A[t] = pa[t];
        B[t] = pb[t];
    }
    __syncthreads();
    // forward stages with flat distance < C
    for (uint32_t m = rows, t = C / 2; m < n2; m *= 2, t /= 2) {
        for (uint32_t idx = threadIdx.x; idx < C / 2; idx += blockDim.x) {
            uint32_t gl = idx / t, jr = idx % t, j = 2 * t * gl + jr;
            uint2 w = wf[m + row * (C / (2 * t)) + gl];
            uint32_t u = A[j], v = mul_shoup(A[j + t], w.x, w.y, p);
            A[j] = add_mod(u, v, p);
            A[j + t] = sub_mod(u, v, p);
            u = B[j];
            v = mul_shoup(B[j + t], w.x, w.y, p);
            B[j] = add_mod(u, v, p);
            B[j + t] = sub_mod(u, v, p);
        }
        __syncthreads();
    }
    // base case: block gb holds A, B mod (X^n1 - zeta_gb), zeta_gb = psi^(2 bitrev(gb) + 1)
    uint32_t res[8];
    {
        uint32_t cnt = 0;
        for (uint32_t idx = threadIdx.x; idx < C; idx += blockDim.x, ++cnt) {
            const uint32_t s = idx % n1, blk = idx / n1, gb = row * (C / n1) + blk;
            const uint32_t* ab = A + blk * n1;
            const uint32_t* bb = B + blk * n1;
            uint32_t lo = 0, hi = 0;
            for (uint32_t t = 0; t <= s; ++t) lo = add_mod(lo, mont(ab[t], bb[s - t], p, pinv), p);
            for (uint32_t t = s + 1; t < n1; ++t) hi = add_mod(hi, mont(ab[t], bb[s + n1 - t], p, pinv), p);
            uint32_t value = lo;
            if (n1 > 1) {
                if (n2 == 1) {
                    value = sub_mod(lo, hi, p);  // zeta = -1
                } else {
                    uint2 w = wf[n2 / 2 + gb / 2];  // zeta_{2g} = w, zeta_{2g+1} = -w
                    uint32_t z = mul_shoup(hi, w.x, w.y, p);
                    value = (gb & 1) ? sub_mod(lo, z, p) : add_mod(lo, z, p);
                }
            }
            res[cnt] = value;
        }
    }
    __syncthreads();
    {
        uint32_t cnt = 0;
        for (uint32_t idx = threadIdx.x; idx < C; idx += blockDim.x, ++cnt) A[idx] = res[cnt];
    }
    __syncthreads();
    // inverse stages with flat distance < C
    for (uint32_t m = n2 / 2, t = n1; m >= rows && m > 0; m /= 2, t *= 2) {
        for (uint32_t idx = threadIdx.x; idx < C / 2; idx += blockDim.x) {
            uint32_t gl = idx / t, jr = idx % t, j = 2 * t * gl + jr;
            uint2 w = wi[m + row * (C / (2 * t)) + gl];
            uint32_t u = A[j], v = A[j + t];
            A[j] = add_mod(u, v, p);
            A[j + t] = mul_shoup(sub_mod(u, v, p), w.x, w.y, p);
        }
        __syncthreads();
    }
    for (uint32_t t = threadIdx.x; t < C; t += blockDim.x) pa[t] = A[t];
}

__global__ __launch_bounds__(256) void ntt_columns_inverse(uint32_t* __restrict__ plane,
                                                           const PrimeConst* __restrict__ consts,
                                                           const uint2* __restrict__ tw, uint32_t n,
                                                           uint32_t n2, uint32_t C, uint32_t rows,
                                                           uint32_t cw) {
    extern __shared__ uint32_t tile[];  // [rows][cw]
    const uint32_t col0 = blockIdx.x * cw, k = blockIdx.y;
    const uint32_t p = consts[k].p;
    uint32_t* data = plane + size_t(k) * n;
    const uint2* w = tw + size_t(k) * n2;
    const uint32_t total = rows * cw;
    for (uint32_t t = threadIdx.x; t < total; t += blockDim.x)
        tile[t] = data[size_t(t / cw) * C + col0 + t % cw];
    __syncthreads();
    for (uint32_t m = rows / 2, half = 1; m >= 1; m /= 2, half *= 2) {
        for (uint32_t t = threadIdx.x; t < total / 2; t += blockDim.x) {
            uint32_t pr = t / cw, c = t % cw;
            uint32_t g = pr / half, jr = pr % half, row = 2 * half * g + jr;
            uint2 tw_g = w[m + g];
            uint32_t u = tile[row * cw + c], v = tile[(row + half) * cw + c];
            tile[row * cw + c] = add_mod(u, v, p);
            tile[(row + half) * cw + c] = mul_shoup(sub_mod(u, v, p), tw_g.x, tw_g.y, p);
        }
        __syncthreads();
    }
    for (uint32_t t = threadIdx.x; t < total; t += blockDim.x)
        data[size_t(t / cw) * C + col0 + t % cw] = tile[t];
}

// ---------------------------------------------------------------------------
// 6. reconstruction: one lane per coefficient, words in [j][N] layout.
// ---------------------------------------------------------------------------
struct ReconstructArgs {
    uint32_t n, limbs, t_words;
    uint32_t s_shift, k_shift;
    uint64_t mu;
};

__global__ __launch_bounds__(64) void reconstruct(const int64_t* __restrict__ words,
                                                  uint32_t* __restrict__ tsc, uint32_t* __restrict__ c,
                                                  const uint32_t* __restrict__ q,
                                                  const uint32_t* __restrict__ neg_p,
                                                  const uint32_t* __restrict__ corr_t,
                                                  const uint32_t* __restrict__ corr_e,
                                                  ReconstructArgs args) {
    __shared__ uint32_t tile[64][65];
    const uint32_t lane = threadIdx.x, i0 = blockIdx.x * 64, i = i0 + lane;
    const uint32_t n = args.n, L = args.limbs, tw = args.t_words;
    const bool valid = i < n;
    const uint32_t ic = valid ? i : 0;  // clamp reads of idle lanes

    // K = floor((E + 2^63) / 2^64), E = sum_j 2^32j words[L+1+j] + corr_e
    uint32_t ew[4];
    {
        int64_t carry = 0;
        for (uint32_t j = 0; j < 4; ++j) {
            int64_t x = (j < 3 ? words[size_t(L + 1 + j) * n + ic] : 0) + int64_t(corr_e[j]) + carry;
            ew[j] = uint32_t(uint64_t(x));
            carry = x >> 32;
        }
    }
    const uint64_t kq = uint64_t(ew[2]) + (uint64_t(ew[3]) << 32) + ((uint64_t(ew[1]) + 0x80000000ull) >> 32);

    // Pass 1: T = sum_j 2^32j words[j] + corr_t + K*R, carried into 32-bit words.
    const uint32_t sw = args.s_shift / 32, sb = args.s_shift % 32;
    uint32_t top[3] = {0, 0, 0};
    {
        int64_t carry = 0;
        for (uint32_t j = 0; j < tw; ++j) {
            int64_t x = (j <= L ? words[size_t(j) * n + ic] : 0) + int64_t(corr_t[j]) + carry;
            if (j < L) x += int64_t(kq) * int64_t(neg_p[j]);
            const uint32_t tj = uint32_t(uint64_t(x));
            carry = x >> 32;
            if (valid) tsc[size_t(j) * n + ic] = tj;
            if (j >= sw && j < sw + 3) top[j - sw] = tj;
        }
    }
    const unsigned __int128 topv = (unsigned __int128)top[0] | ((unsigned __int128)top[1] << 32) |
                                   ((unsigned __int128)top[2] << 64);
    const uint64_t ts = uint64_t(topv >> sb);
    const uint64_t qe = uint64_t(((unsigned __int128)ts * args.mu) >> args.k_shift);

    // Pass 2: rem = T - qe*q (in [0, 2q)), emitted through an LDS transpose;
    // a second borrow chain tracks rem - q.
    uint64_t pc = 0;
    uint32_t bo = 0, bo2 = 0;
    for (uint32_t j0 = 0; j0 < tw; j0 += 64) {
        const uint32_t jend = j0 + 64 < tw ? j0 + 64 : tw;
        for (uint32_t j = j0; j < jend; ++j) {
            const uint32_t qj = j < L ? q[j] : 0;
            const unsigned __int128 pr = (unsigned __int128)qe * qj + pc;
            const uint32_t x = uint32_t(pr);
            pc = uint64_t(pr >> 32);
            const uint32_t tj = tsc[size_t(j) * n + ic];
            const uint64_t d = uint64_t(tj) - x - bo;
            const uint32_t rj = uint32_t(d);
            bo = uint32_t(d >> 63);
            const uint64_t d2 = uint64_t(rj) - qj - bo2;
            bo2 = uint32_t(d2 >> 63);
            tile[lane][j - j0] = rj;
        }
        __syncthreads();
        // row-wise coalesced store: c[i0+row][j0 .. jend) for j < L
        for (uint32_t row = 0; row < 64; ++row) {
            const uint32_t gi = i0 + row, j = j0 + lane;
            if (gi < n && j < jend && j < L) c[size_t(gi) * L + j] = tile[row][lane];
        }
        __syncthreads();
    }
    // rem >= q (the estimate was one short): subtract q once more, in place.
    __syncthreads();
    if (valid && bo2 == 0) {
        uint32_t borrow = 0;
        uint32_t* ci = c + size_t(i) * L;
        for (uint32_t j = 0; j < L; ++j) {
            const uint64_t d = uint64_t(ci[j]) - q[j] - borrow;
            ci[j] = uint32_t(d);
            borrow = uint32_t(d >> 63);
        }
    }
}

__global__ void fill_zero(uint32_t* __restrict__ c, size_t words) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < words) c[i] = 0;
}

uint32_t column_width(const Shape& s) {
    uint32_t cw = 8192 / s.rows;
    if (cw > s.row_width) cw = s.row_width;
    if (cw < 1) cw = 1;
    return cw;
}

}  // namespace

void launch_conversion(const Shape& s, const DeviceTables& t, const uint32_t* a, const uint32_t* b,
                       uint32_t* planes, hipStream_t stream) {
    dim3 grid((s.n + kBN - 1) / kBN, s.conv_rows / kBM, 2);
    hipLaunchKernelGGL(conversion_gemm, grid, dim3(kGemmThreads), 0, stream, t.wconv, a, b, t.consts, planes,
                       s.n, s.limbs, s.r, s.conv_rows, s.conv_k);
    check_launch("conversion_gemm");
}

void launch_ntt_forward_columns(const Shape& s, const DeviceTables& t, uint32_t* planes, hipStream_t stream) {
    if (s.rows < 2) return;
    const uint32_t cw = column_width(s);
    dim3 grid(s.row_width / cw, s.r, 2);
    hipLaunchKernelGGL(ntt_columns_forward, grid, dim3(256), size_t(s.rows) * cw * 4, stream, planes, t.consts,
                       t.tw_fwd, s.n, s.n2, s.row_width, s.rows, cw, s.r);
    check_launch("ntt_columns_forward");
}

void launch_ntt_rows_product(const Shape& s, const DeviceTables& t, uint32_t* planes, hipStream_t stream) {
    dim3 grid(s.rows, s.r);
    hipLaunchKernelGGL(ntt_rows_product, grid, dim3(256), size_t(2) * s.row_width * 4, stream, planes, t.consts,
                       t.tw_fwd, t.tw_inv, s.n, s.n1, s.n2, s.row_width, s.rows, s.r);
    check_launch("ntt_rows_product");
}

void launch_ntt_inverse_columns(const Shape& s, const DeviceTables& t, uint32_t* planes, hipStream_t stream) {
    if (s.rows < 2) return;
    const uint32_t cw = column_width(s);
    dim3 grid(s.row_width / cw, s.r);
    hipLaunchKernelGGL(ntt_columns_inverse, grid, dim3(256), size_t(s.rows) * cw * 4, stream, planes, t.consts,
                       t.tw_inv, s.n, s.n2, s.row_width, s.rows, cw);
    check_launch("ntt_columns_inverse");
}

void launch_crt(const Shape& s, const DeviceTables& t, const uint32_t* product_plane, int64_t* words,
                hipStream_t stream) {
    dim3 grid((s.n + kBN - 1) / kBN, s.crt_rows / kBM);
    hipLaunchKernelGGL(crt_gemm, grid, dim3(kGemmThreads), 0, stream, t.wcrt, product_plane, words, s.n, s.r,
                       s.crt_k, s.crt_words);
    check_launch("crt_gemm");
}

void launch_reconstruct(const Shape& s, const DeviceTables& t, const int64_t* words, uint32_t* t_scratch,
                        uint32_t* c, hipStream_t stream) {
    ReconstructArgs args{s.n, s.limbs, s.t_words, s.s_shift, s.k_shift, s.mu};
    dim3 grid((s.n + 63) / 64);
    hipLaunchKernelGGL(reconstruct, grid, dim3(64), 0, stream, words, t_scratch, c, t.q, t.neg_p, t.corr_t,
                       t.corr_e, args);
    check_launch("reconstruct");
}

void launch_zero(uint32_t* c, size_t words, hipStream_t stream) {
    if (!words) return;
    dim3 grid(uint32_t((words + 255) / 256));
    hipLaunchKernelGGL(fill_zero, grid, dim3(256), 0, stream, c, words);
    check_launch("fill_zero");
}

}  // namespace pm
