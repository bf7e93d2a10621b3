// Internal declarations shared by the host C++ and HIP translation units of
// libqvnt_hip.so.  The public contract is include/qvnt.h.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/qvnt.h"

struct qvnt_op {
    std::vector<qvnt_gate> g;  // reference MultiOp: VecDeque<SingleOp>
};

namespace qvnt {

// thread-local last error (qvnt_last_error)
int fail(int code, const std::string& msg);
void clear_error();

inline unsigned popc(uint64_t x) { return (unsigned)__builtin_popcountll(x); }
inline unsigned ctz(uint64_t x) { return (unsigned)__builtin_ctzll(x); }
inline unsigned msb(uint64_t x) { return 63u - (unsigned)__builtin_clzll(x); }

// Gate classes, by the shape of the index orbit one thread owns.
enum class GClass { Noop, Diag, Pair, Quad };

// Bits a gate reads besides idx itself (the orbit's span).
inline uint64_t touched_bits(const qvnt_gate& g) {
    switch (g.kind) {
    case QVNT_H2:
    case QVNT_U2: return g.a | g.b;
    default: return g.a;
    }
}

bool is_diagonal_kind(uint32_t kind);
// the reference's unitarity rule for a 2x2 complex matrix (interleaved re, im)
bool is_unitary_m1(const double* m);
// num-complex's Debug of a Complex<f64> ("Complex { re: 0.25, im: 0.0 }")
std::string complex_debug_value(double re, double im);

}  // namespace qvnt
