// Device-side building blocks shared by the HIP translation units:
// complex amplitudes, index deposit, and the per-gate formulas.
//
// The formulas restate the reference's AtomicOp::atomic_op of every gate
// (src/operator/atomic/<gate>.rs) expression by expression, so that with
// FP contraction off a double-precision run rounds exactly like the
// reference's CPU loop.  num-complex's Mul is (a.re*b.re - a.im*b.im,
// a.re*b.im + a.im*b.re) with `a` the left operand.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/qvnt.h"

namespace qvnt {

template <class R>
struct alignas(2 * sizeof(R)) cplx {
    R re, im;
};
using c64 = cplx<double>;
using c32 = cplx<float>;

template <class R>
__device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class R>
__device__ __forceinline__ cplx<R> cadd(cplx<R> a, cplx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <class R>
__device__ __forceinline__ cplx<R> cneg(cplx<R> a) { return {-a.re, -a.im}; }
template <class R>
__device__ __forceinline__ cplx<R> cscale(cplx<R> a, R t) { return {a.re * t, a.im * t}; }
// math::rotate (src/math/mod.rs:41-50): i^q * z, q mod 4.
template <class R>
__device__ __forceinline__ cplx<R> rotate(cplx<R> z, uint64_t q) {
    if (q & 2u) z = cneg(z);
    if (q & 1u) z = cplx<R>{-z.im, z.re};
    return z;
}
__device__ __forceinline__ unsigned dpopc(uint64_t x) { return (unsigned)__popcll(x); }

// Bit positions inserted into a compact thread index to enumerate one
// representative per orbit: `pos` ascending, then `set` ORed in (control
// bits, and a fixed 1 where the orbit needs one).
// Fixed bits of a launch: `zeros` holds their positions, `set` the ones
// among them that are 1, n = popcount(zeros).  The walk over `zeros` is
// uniform scalar work (no table loads in the kernels' address path).
// `skew` (kernels with two units per thread): 0, or S >= 1 for a thread
// whose two units are 2^S apart in the compact index (see unit_index).
struct Deposit {
    uint64_t zeros;
    uint64_t set;
    uint32_t n;
    uint32_t skew;
};

// Compact index of unit u (0 .. 2^ubits - 1) of thread tau when the units
// sit 2^S apart: bits S .. S + ubits - 1 of the index are u, the bits of tau
// close around them.
__device__ __forceinline__ uint64_t unit_index(uint64_t tau, uint32_t u, uint32_t S, uint32_t ubits = 1) {
    return ((tau >> S) << (S + ubits)) | ((uint64_t)u << S) | (tau & ((1ull << S) - 1u));
}

__device__ __forceinline__ uint64_t deposit(uint64_t g, const Deposit& d) {
    for (uint64_t m = d.zeros; m; m &= m - 1) {
        const uint64_t lo = g & ((m & (0 - m)) - 1u);  // bits below the lowest position left
        g = ((g ^ lo) << 1) | lo;
    }
    return g | d.set;
}

// Gate parameters, FP64 for either amplitude precision: the stored phases
// (the QFT ladder's controlled-RZ angles among them) and matrices keep the
// reference's double values, and the formulas below compute in FP64.
struct Par {
    uint64_t a, b;       // a: act mask (orbit mask for pair kinds); b: H2/U2 b_mask
    double re, im;       // stored phase
    double m[32];        // U1 / U2 matrices
    uint32_t dagger;
    uint32_t i_pow;      // Y: !(popcount(a) + 1) as usize (atomic/y.rs:11)
    // Index bits tested per amplitude instead of being fixed in the
    // enumeration (bits inside one 128-byte line, so that every line is read
    // and written whole): the gate acts where (idx & pred) == pred (its
    // controls there) and, for the swap family, where idx & podd has odd
    // parity; elsewhere the amplitude is written back unchanged, which is
    // the reference's copy of psi_i[idx] (dispatch.rs:35-39).
    uint64_t pred, podd;
};

// Whether the gate acts on the orbit of idx (see Par::pred / Par::podd).
__device__ __forceinline__ bool acts(const Par& g, uint64_t idx) {
    return (idx & g.pred) == g.pred && (g.podd == 0 || (dpopc(idx & g.podd) & 1u));
}

constexpr double kFrac1Sqrt2 = 0.70710678118654752440;

// Amplitudes enter a formula in FP64 and leave it rounded once to the
// register's precision R (for R = double both conversions are the identity,
// so FP64 registers round exactly like the reference).
template <class R>
__device__ __forceinline__ c64 up(cplx<R> p) { return c64{(double)p.re, (double)p.im}; }
template <class R>
__device__ __forceinline__ cplx<R> down(c64 p) { return cplx<R>{(R)p.re, (R)p.im}; }

// ------------------------------------------------------------- diagonals
// out[idx] = f(psi[idx]); kinds Z, S, T, RZ, RZZ, PHI.
template <int K>
__device__ __forceinline__ c64 f_diag64(const Par& g, uint64_t idx, c64 p) {
    if constexpr (K == QVNT_Z) {                                   // z.rs:15-21
        return (dpopc(idx & g.a) & 1u) ? cneg(p) : p;
    } else if constexpr (K == QVNT_S) {                            // s.rs:19-25
        uint64_t count = dpopc(idx & g.a);
        if (g.dagger) count = (~count) + 1u;
        return rotate(p, count);
    } else if constexpr (K == QVNT_T) {                            // t.rs:24-35
        uint64_t count = dpopc(idx & g.a);
        if (g.dagger) count = (~count) + 1u;
        const c64 z = rotate(p, count >> 1);
        const c64 e{kFrac1Sqrt2, kFrac1Sqrt2};
        return (count & 1u) ? cmul(e, z) : z;
    } else if constexpr (K == QVNT_RZ) {                           // rz.rs:19-25
        c64 ph{g.re, g.im};
        if ((idx & g.a) == 0) ph.im = -ph.im;
        return cmul(ph, p);
    } else if constexpr (K == QVNT_RZZ) {                          // rzz.rs:19-25
        c64 ph{g.re, g.im};
        if ((dpopc(idx & g.a) & 1u) == 0) ph.im = -ph.im;
        return cmul(ph, p);
    } else {                                                       // PHI: diag(1, e^{i th})
        const c64 ph{g.re, g.im};
        return (idx & g.a) ? cmul(ph, p) : p;
    }
}
template <int K, class R>
__device__ __forceinline__ cplx<R> f_diag(const Par& g, uint64_t idx, cplx<R> p) {
    return down<R>(f_diag64<K>(g, idx, up(p)));
}

// ---------------------------------------------------------- pair orbits
// out[idx] = f(psi[idx], psi[idx ^ m]); `self` = psi[idx],
// `other` = psi[idx ^ m] with m the orbit mask.
template <int K>
__device__ __forceinline__ c64 f_pair64(const Par& g, uint64_t idx, c64 self, c64 other) {
    if constexpr (K == QVNT_X) {                                   // x.rs:15-17
        return other;
    } else if constexpr (K == QVNT_Y) {                            // y.rs:17-23
        uint64_t q = g.i_pow;
        if ((dpopc(idx & g.a) & 1u) == 0) q ^= 2u;
        return rotate(other, q);
    } else if constexpr (K == QVNT_RX || K == QVNT_RXX) {          // rx.rs:18-24, rxx.rs:19-25
        return {self.re * g.re + other.im * g.im, self.im * g.re - other.re * g.im};
    } else if constexpr (K == QVNT_RY) {                           // ry.rs:19-29
        double s = g.im;
        if ((idx & g.a) == 0) s = -s;
        return {self.re * g.re + other.re * s, self.im * g.re + other.im * s};
    } else if constexpr (K == QVNT_RYY) {                          // ryy.rs:19-29
        double s = g.im;
        if ((dpopc(idx & g.a) & 1u) == 0) s = -s;
        return {self.re * g.re + other.im * s, self.im * g.re - other.re * s};
    } else if constexpr (K == QVNT_H1) {                           // h1.rs:16-22
        c64 p0 = self;
        if (idx & g.a) p0 = cneg(p0);
        return cscale(cadd(p0, other), kFrac1Sqrt2);
    } else if constexpr (K == QVNT_U1) {                           // u1.rs:17-25
        if ((idx & g.a) == 0)
            return cadd(cmul(c64{g.m[0], g.m[1]}, self), cmul(c64{g.m[2], g.m[3]}, other));
        return cadd(cmul(c64{g.m[4], g.m[5]}, other), cmul(c64{g.m[6], g.m[7]}, self));
    } else if constexpr (K == QVNT_SWAP) {                         // swap.rs:16-22 (odd orbits)
        return other;
    } else if constexpr (K == QVNT_ISWAP) {                        // i_swap.rs:20-37 (odd orbits)
        if (g.dagger) return {other.im, -other.re};
        return {-other.im, other.re};
    } else if constexpr (K == QVNT_SQRT_SWAP) {                    // sqrt_swap.rs:20-37 (odd orbits)
        const c64 p0 = self, p1 = other;
        if (g.dagger)
            return {0.5 * (p0.re + p0.im + p1.re - p1.im), 0.5 * (p0.im - p0.re + p1.im + p1.re)};
        return {0.5 * (p0.re - p0.im + p1.re + p1.im), 0.5 * (p0.im + p0.re + p1.im - p1.re)};
    } else {                                                       // sqrt_i_swap.rs:20-37 (odd orbits)
        const c64 p0 = self, p1 = other;
        if (g.dagger) return {kFrac1Sqrt2 * (p0.re + p1.im), kFrac1Sqrt2 * (p0.im - p1.re)};
        return {kFrac1Sqrt2 * (p0.re - p1.im), kFrac1Sqrt2 * (p0.im + p1.re)};
    }
}
template <int K, class R>
__device__ __forceinline__ cplx<R> f_pair(const Par& g, uint64_t idx, cplx<R> self, cplx<R> other) {
    return down<R>(f_pair64<K>(g, idx, up(self), up(other)));
}

// ---------------------------------------------------------- quad orbits
// v[k] = psi[base | sel_k] with sel bits (b, a) = k; output for v-slot `s`.
// h2.rs:22-38 for slot s, from the quad as seen from s: t0 = v[s],
// t1 = v[s ^ 1], t2 = v[s ^ 2], t3 = v[s ^ 3].
__device__ __forceinline__ c64 f_h2_rel(int s, c64 t0, c64 t1, c64 t2, c64 t3) {
    if (s & 1) { t0 = cneg(t0); t2 = cneg(t2); }
    if (s & 2) { t0 = cneg(t0); t1 = cneg(t1); }
    return cscale(cadd(cadd(cadd(t0, t1), t2), t3), 0.5);
}

template <int K>
__device__ __forceinline__ c64 f_quad64(const Par& g, int s, const c64* v) {
    if constexpr (K == QVNT_H2) {                                  // h2.rs:22-38
        return f_h2_rel(s, v[s], v[s ^ 1], v[s ^ 2], v[s ^ 3]);
    } else {                                                       // u2.rs:22-49
        const double* m = g.m + 8 * s;
        c64 acc = cmul(c64{m[0], m[1]}, v[0]);
        acc = cadd(acc, cmul(c64{m[2], m[3]}, v[1]));
        acc = cadd(acc, cmul(c64{m[4], m[5]}, v[2]));
        acc = cadd(acc, cmul(c64{m[6], m[7]}, v[3]));
        return acc;
    }
}
template <int K, class R>
__device__ __forceinline__ cplx<R> f_quad(const Par& g, int s, const cplx<R>* v) {
    const c64 w[4] = {up(v[0]), up(v[1]), up(v[2]), up(v[3])};
    return down<R>(f_quad64<K>(g, s, w));
}

}  // namespace qvnt
