// Drop-in C++ surface of the FHERMA polynomial-multiplication/negacyclic@1.0.0
// harness, implemented on MI355X by libfherma_amd.so over the C ABI of
// include/polymul.h.
//
// The types and the three entry points are the harness's interface
// (reference: src/fherma.h:13-48), so they are reproduced with the same names,
// layout and meaning: a Point carries N, W (bits of q), L = ceil(W/32) and q
// as L little-endian 32-bit limbs; operands and the result are row-major
// [N][L] limb tensors of canonical coefficients in [0, q).  Errors surface as
// std::runtime_error, as the harness expects.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace fherma {

template <class T>
struct Tensor {
    std::vector<int64_t> shape;
    std::vector<T> data;  // row-major

    int64_t count() const {
        int64_t total = 1;
        for (auto d : shape) total *= d;
        return total;
    }
};

struct Point {
    uint32_t N = 0;        // ring degree, a power of two
    uint32_t W = 0;        // bits of q
    uint32_t L = 0;        // ceil(W / 32) limbs per coefficient
    Tensor<uint32_t> q;    // [L]
};

struct Inputs {
    Tensor<uint32_t> a;    // [N][L]
    Tensor<uint32_t> b;    // [N][L]
};

struct Outputs {
    Tensor<uint32_t> c;    // [N][L]
};

}  // namespace fherma

// Parameter-only initialization (replaces src/fherma.h:46).
void* fherma_init(const fherma::Point& p);
// One product, host tensors in and out (replaces src/fherma.h:47).
fherma::Outputs fherma_run(void* state, const fherma::Inputs& in);
// Releases the state (replaces src/fherma.h:48).
void fherma_free(void* state);
