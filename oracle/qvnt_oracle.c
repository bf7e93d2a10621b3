/*
 * qvnt_oracle.c — CPU restatement of qvnt 0.4.4 (TEST INFRASTRUCTURE ONLY:
 * the parity oracle and the CPU baseline; never part of the product path).
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 * Paths cited are relative to the reference root.
 */
#define _GNU_SOURCE
#include "qvnt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define FRAC_1_SQRT_2 0.70710678118654752440 /* std::f64::consts::FRAC_1_SQRT_2 */
#define PI_F64 3.14159265358979323846        /* std::f64::consts::PI            */
#define FRAC_PI_2_F64 1.57079632679489661923 /* std::f64::consts::FRAC_PI_2     */
#define MIN_BUFFER_LEN 8u                    /* src/register/quant.rs:15        */

/* ------------------------------------------------------------ complex */
/* num-complex 0.4 arithmetic (Mul/Add/Neg/scale), in its operand order. */
static inline orc_c c_mul(orc_c a, orc_c b) {
    orc_c r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline orc_c c_add(orc_c a, orc_c b) { orc_c r = {a.re + b.re, a.im + b.im}; return r; }
static inline orc_c c_neg(orc_c a) { orc_c r = {-a.re, -a.im}; return r; }
static inline orc_c c_scale(orc_c a, double t) { orc_c r = {a.re * t, a.im * t}; return r; }
static inline orc_c c_conj(orc_c a) { orc_c r = {a.re, -a.im}; return r; }
static inline double c_norm_sqr(orc_c a) { return a.re * a.re + a.im * a.im; }

/* math::rotate (src/math/mod.rs:41-50): multiply by i^q, q mod 4. */
static inline orc_c rotate(orc_c z, uint64_t q) {
    if (q & 2u) z = c_neg(z);
    if (q & 1u) {
        double t;
        z.im = -z.im;
        t = z.re; z.re = z.im; z.im = t;
    }
    return z;
}

static inline unsigned popc(uint64_t x) { return (unsigned)__builtin_popcountll(x); }

/* ------------------------------------------------------------ builders */
static void gate_init(qvnt_gate* g, uint32_t kind, uint64_t act) {
    memset(g, 0, sizeof(*g));
    g->kind = kind;
    g->act = act;
    g->a = act;
}

/* Rotation constructors store phase = (cos(th/2), sin(th/2))
 * (atomic/rx.rs:10-14, ry.rs:11-15, rz.rs:11-15, rzz.rs:11-15, ryy.rs:11-15;
 * rxx.rs:11-15 uses `*= 0.5`, the same value). */
static long rot(uint32_t kind, double phase, uint64_t mask, unsigned bits, qvnt_gate* out) {
    if (popc(mask) != bits) return QVNT_E_INVALID_MASK;  /* is_valid + expect */
    gate_init(out, kind, mask);
    phase = (kind == QVNT_RXX) ? phase * 0.5 : phase / 2.0;
    /* C::new(phase.cos(), phase.sin()): an optimised Rust build on
     * linux-gnu lowers the cos/sin pair of one argument to glibc sincos. */
    sincos(phase, &out->im, &out->re);
    return 1;
}

long orc_op_id(qvnt_gate* out) { (void)out; return 0; }  /* operator/mod.rs:113-115: empty */
/* op::x/y/z/s/t (mod.rs:136,199,258,277,297 -> single/pauli.rs:7-30): any mask. */
long orc_op_x(uint64_t a, qvnt_gate* out) { gate_init(out, QVNT_X, a); return 1; }
long orc_op_y(uint64_t a, qvnt_gate* out) { gate_init(out, QVNT_Y, a); return 1; }
long orc_op_z(uint64_t a, qvnt_gate* out) { gate_init(out, QVNT_Z, a); return 1; }
long orc_op_s(uint64_t a, qvnt_gate* out) { gate_init(out, QVNT_S, a); return 1; }
long orc_op_t(uint64_t a, qvnt_gate* out) { gate_init(out, QVNT_T, a); return 1; }
/* op::rx/ry/rz/rxx/ryy/rzz (mod.rs:151,214,321,170,233,340 -> single/rotate.rs). */
long orc_op_rx(double p, uint64_t a, qvnt_gate* out) { return rot(QVNT_RX, p, a, 1, out); }
long orc_op_ry(double p, uint64_t a, qvnt_gate* out) { return rot(QVNT_RY, p, a, 1, out); }
long orc_op_rz(double p, uint64_t a, qvnt_gate* out) { return rot(QVNT_RZ, p, a, 1, out); }
long orc_op_rxx(double p, uint64_t ab, qvnt_gate* out) { return rot(QVNT_RXX, p, ab, 2, out); }
long orc_op_ryy(double p, uint64_t ab, qvnt_gate* out) { return rot(QVNT_RYY, p, ab, 2, out); }
long orc_op_rzz(double p, uint64_t ab, qvnt_gate* out) { return rot(QVNT_RZZ, p, ab, 2, out); }
/* phi: listed in README.md:69, absent from the code; defined here as
 * diag(1, e^{i th}) on one qubit, stored as (cos th, sin th). */
long orc_op_phi(double p, uint64_t a, qvnt_gate* out) {
    if (popc(a) != 1) return QVNT_E_INVALID_MASK;
    gate_init(out, QVNT_PHI, a);
    sincos(p, &out->im, &out->re);
    return 1;
}
/* op::swap family (mod.rs:365-456 -> single/swap.rs:7-24): exactly 2 bits. */
static long swp(uint32_t kind, uint64_t ab, qvnt_gate* out) {
    if (popc(ab) != 2) return QVNT_E_INVALID_MASK;
    gate_init(out, kind, ab);
    return 1;
}
long orc_op_swap(uint64_t ab, qvnt_gate* out) { return swp(QVNT_SWAP, ab, out); }
long orc_op_sqrt_swap(uint64_t ab, qvnt_gate* out) { return swp(QVNT_SQRT_SWAP, ab, out); }
long orc_op_i_swap(uint64_t ab, qvnt_gate* out) { return swp(QVNT_ISWAP, ab, out); }
long orc_op_sqrt_i_swap(uint64_t ab, qvnt_gate* out) { return swp(QVNT_SQRT_ISWAP, ab, out); }

static void h1(uint64_t a, qvnt_gate* out) { gate_init(out, QVNT_H1, a); }
/* atomic/h2.rs:12-19: ab = a | b; acts_on = ab. */
static void h2(uint64_t a, uint64_t b, qvnt_gate* out) {
    gate_init(out, QVNT_H2, a | b);
    out->a = a;
    out->b = b;
}

/* op::h -> multi/h.rs:14-45: consecutive set bits paired into H2(higher,
 * lower); an odd bit left over becomes H1. */
long orc_op_h(uint64_t a, qvnt_gate* out) {
    unsigned count = popc(a);
    long n = 0;
    if (count == 0) return 0;
    if (count == 1) { h1(a, out); return 1; }
    {
        uint64_t first = 0;
        int is_first = 1;
        unsigned k;
        for (k = 0; k < 64; ++k) {
            uint64_t bit = 1ull << k;
            if (bit > a) break;
            if (bit & a) {
                if (is_first) { first = bit; is_first = 0; }
                else { h2(bit, first, &out[n++]); is_first = 1; }
            }
        }
        if (!is_first) h1(first, &out[n++]);
    }
    return n;
}

/* op::u1 = rz (mod.rs:472-474); u2 = rz(lam)*ry(pi/2)*rz(phi) (mod.rs:480-482);
 * u3 = rz(lam)*ry(the)*rz(phi) (mod.rs:499-501); `*` appends (multi/mod.rs:166). */
long orc_op_u1(double lam, uint64_t a, qvnt_gate* out) { return orc_op_rz(lam, a, out); }
long orc_op_u2(double phi, double lam, uint64_t a, qvnt_gate* out) {
    long r;
    if ((r = orc_op_rz(lam, a, &out[0])) < 0) return r;
    if ((r = orc_op_ry(FRAC_PI_2_F64, a, &out[1])) < 0) return r;
    if ((r = orc_op_rz(phi, a, &out[2])) < 0) return r;
    return 3;
}
long orc_op_u3(double the, double phi, double lam, uint64_t a, qvnt_gate* out) {
    long r;
    if ((r = orc_op_rz(lam, a, &out[0])) < 0) return r;
    if ((r = orc_op_ry(the, a, &out[1])) < 0) return r;
    if ((r = orc_op_rz(phi, a, &out[2])) < 0) return r;
    return 3;
}

/* op::qft -> multi/qft.rs:4-33: for i < count-1: h(bit_i), then
 * rz(bit_{i+j}, pi * 0.5^j).c(bit_i) for j = 1 .. count-i-1; finally
 * h(bit_{count-1}). */
long orc_op_qft(uint64_t a, qvnt_gate* out) {
    unsigned count = popc(a), i, j, k, m = 0;
    uint64_t vec[64];
    long n = 0;
    if (count == 0) return 0;
    if (count == 1) return orc_op_h(a, out);
    for (k = 0; k < 64; ++k)
        if (a & (1ull << k)) vec[m++] = 1ull << k;
    for (i = 0; i + 1 < count; ++i) {
        n += orc_op_h(vec[i], &out[n]);
        for (j = 1; j < count - i; ++j) {
            double p = PI_F64 * ldexp(1.0, -(int)j); /* 0.5f64.powi(j): exact */
            orc_op_rz(p, vec[i + j], &out[n]);
            out[n].ctrl |= vec[i];                   /* .c(vec[i]).unwrap() */
            ++n;
        }
    }
    n += orc_op_h(vec[count - 1], &out[n]);
    return n;
}

/* op::qft_swapped -> multi/qft.rs:35-52: qft then swaps of mirrored bits. */
long orc_op_qft_swapped(uint64_t a, qvnt_gate* out) {
    uint64_t vec[64];
    unsigned len = 0, k, i;
    long n = orc_op_qft(a, out);
    for (k = 0; k < 64; ++k)
        if (a & (1ull << k)) vec[len++] = 1ull << k;
    for (i = 0; i < (len >> 1); ++i) n += orc_op_swap(vec[i] | vec[len - i - 1], &out[n]);
    return n;
}

/* float-cmp approx_eq!(f64, x, y, ulps = 2) with the default epsilon
 * (math/approx_cmp.rs:5-8): |x - y| <= f64::EPSILON or within 2 ulps. */
static int approx_eq_real(double x, double y) {
    int64_t ix, iy, d;
    if (fabs(x - y) <= 2.220446049250313e-16) return 1;
    if ((x < 0) != (y < 0)) return x == y;
    memcpy(&ix, &x, 8);
    memcpy(&iy, &y, 8);
    d = ix > iy ? ix - iy : iy - ix;
    return d <= 2;
}
static orc_c mget(const double* m, int k) { orc_c c = {m[2 * k], m[2 * k + 1]}; return c; }

/* is_unitary_m1 (math/matrix.rs:23-28). */
static int is_unitary_m1(const double* m) {
    orc_c u00 = mget(m, 0), u01 = mget(m, 1), u10 = mget(m, 2), u11 = mget(m, 3);
    double e00 = c_norm_sqr(u00) + c_norm_sqr(u01);
    double e11 = c_norm_sqr(u10) + c_norm_sqr(u11);
    orc_c e01 = c_add(c_mul(u00, c_conj(u10)), c_mul(u01, c_conj(u11)));
    return approx_eq_real(e00, 1.0) && approx_eq_real(e11, 1.0) && approx_eq_real(e01.re + e01.im, 0.0);
}
/* hermitian_mul / is_unitary_m2 (math/matrix.rs:35-66). */
static orc_c herm(int i, int j, const double* m) {
    i = (i << 2) & 0xf;
    j = (j << 2) & 0xf;
    if (i == j) {
        orc_c r = {(c_norm_sqr(mget(m, 0 | i)) + c_norm_sqr(mget(m, 1 | i))) +
                       (c_norm_sqr(mget(m, 2 | i)) + c_norm_sqr(mget(m, 3 | i))),
                   0.0};
        return r;
    }
    return c_add(c_add(c_mul(mget(m, 0 | i), c_conj(mget(m, 0 | j))), c_mul(mget(m, 1 | i), c_conj(mget(m, 1 | j)))),
                 c_add(c_mul(mget(m, 2 | i), c_conj(mget(m, 2 | j))), c_mul(mget(m, 3 | i), c_conj(mget(m, 3 | j)))));
}
static int is_unitary_m2(const double* m) {
    int i, j;
    for (i = 0; i < 4; ++i)
        if (!approx_eq_real(herm(i, i, m).re, 1.0)) return 0;
    for (i = 0; i < 4; ++i)
        for (j = i + 1; j < 4; ++j) {
            orc_c e = herm(i, j, m);
            if (!approx_eq_real(e.re + e.im, 0.0)) return 0;
        }
    return 1;
}
/* single/pauli.rs:32-39 (u1/u2 matrix constructors, checked). */
long orc_op_matrix1(uint64_t a, const double m[8], qvnt_gate* out) {
    if (popc(a) != 1 || !is_unitary_m1(m)) return QVNT_E_INVALID_MASK;
    gate_init(out, QVNT_U1, a);
    memcpy(out->m, m, 8 * sizeof(double));
    return 1;
}
long orc_op_matrix2(uint64_t a, uint64_t b, const double m[32], qvnt_gate* out) {
    if (popc(a) != 1 || popc(b) != 1 || !is_unitary_m2(m)) return QVNT_E_INVALID_MASK;
    gate_init(out, QVNT_U2, a);  /* acts_on() returns only a_mask (atomic/u2.rs:70-72) */
    out->a = a;
    out->b = b;
    memcpy(out->m, m, 32 * sizeof(double));
    return 1;
}

/* SingleOp::act_on = act | ctrl (single/mod.rs:96-98); MultiOp folds it
 * (multi/mod.rs:116-118). */
uint64_t orc_op_act_on(const qvnt_gate* g, size_t n) {
    uint64_t r = 0;
    size_t i;
    for (i = 0; i < n; ++i) r |= g[i].act | g[i].ctrl;
    return r;
}
/* MultiOp::c (multi/mod.rs:125-132). */
int orc_op_c(qvnt_gate* g, size_t n, uint64_t c) {
    size_t i;
    if (orc_op_act_on(g, n) & c) return QVNT_E_CTRL_OVERLAP;
    for (i = 0; i < n; ++i) g[i].ctrl |= c;
    return 0;
}
/* AtomicOp::dgr of each kind (atomic/<gate>.rs `fn dgr`). */
static void gate_dgr(qvnt_gate* g) {
    switch (g->kind) {
    case QVNT_S: case QVNT_T: case QVNT_ISWAP: case QVNT_SQRT_SWAP: case QVNT_SQRT_ISWAP:
        g->dagger = !g->dagger;  /* s.rs:42-47, t.rs:52-57, i_swap.rs:58-63, ... */
        break;
    case QVNT_RX: case QVNT_RY: case QVNT_RZ: case QVNT_RXX: case QVNT_RYY: case QVNT_RZZ:
        g->re = -g->re;          /* `phase: -self.phase` (rx.rs:44 etc.): -U */
        g->im = -g->im;
        break;
    case QVNT_PHI:
        g->im = -g->im;          /* diag(1, e^{-i th}) */
        break;
    case QVNT_U1: {              /* inverse_unitary_m1 (math/matrix.rs:30-33) */
        double t[8];
        int r, c;
        for (r = 0; r < 2; ++r)
            for (c = 0; c < 2; ++c) {
                t[2 * (2 * r + c)] = g->m[2 * (2 * c + r)];
                t[2 * (2 * r + c) + 1] = -g->m[2 * (2 * c + r) + 1];
            }
        memcpy(g->m, t, sizeof(t));
        break;
    }
    case QVNT_U2: {              /* inverse_unitary_m2 (math/matrix.rs:68-88) */
        double t[32];
        int r, c;
        for (r = 0; r < 4; ++r)
            for (c = 0; c < 4; ++c) {
                t[2 * (4 * r + c)] = g->m[2 * (4 * c + r)];
                t[2 * (4 * r + c) + 1] = -g->m[2 * (4 * c + r) + 1];
            }
        memcpy(g->m, t, sizeof(t));
        break;
    }
    default:
        break;                   /* X, Y, Z, H1, H2, SWAP, Id: self */
    }
}
/* MultiOp::dgr: dgr each, reverse (multi/mod.rs:120-123). */
void orc_op_dgr(qvnt_gate* g, size_t n) {
    size_t i;
    for (i = 0; i < n; ++i) gate_dgr(&g[i]);
    for (i = 0; i < n / 2; ++i) {
        qvnt_gate t = g[i];
        g[i] = g[n - 1 - i];
        g[n - 1 - i] = t;
    }
}

/* ------------------------------------------------- the per-gate gather */
/* AtomicOp::atomic_op(psi, idx) of every kind (src/operator/atomic/<gate>.rs). */
static inline orc_c atomic_op(const qvnt_gate* g, const orc_c* psi, uint64_t idx) {
    const uint64_t a = g->a;
    switch (g->kind) {
    case QVNT_ID:                                              /* id.rs:7-9 */
        return psi[idx];
    case QVNT_X:                                               /* x.rs:15-17 */
        return psi[idx ^ a];
    case QVNT_Y: {                                             /* y.rs:11,17-23 */
        uint64_t i_pow = (uint64_t)(uint32_t)(~(uint32_t)(popc(a) + 1u));
        if ((popc(idx & a) & 1u) == 0) i_pow ^= 2u;
        return rotate(psi[idx ^ a], i_pow);
    }
    case QVNT_Z:                                               /* z.rs:15-21 */
        return (popc(idx & a) & 1u) ? c_neg(psi[idx]) : psi[idx];
    case QVNT_S: {                                             /* s.rs:19-25 */
        uint64_t count = popc(idx & a);
        if (g->dagger) count = (~count) + 1u;
        return rotate(psi[idx], count);
    }
    case QVNT_T: {                                             /* t.rs:3-6,24-35 */
        static const orc_c EXP_I_PI_4 = {FRAC_1_SQRT_2, FRAC_1_SQRT_2};
        uint64_t count = popc(idx & a);
        orc_c z;
        if (g->dagger) count = (~count) + 1u;
        z = rotate(psi[idx], count >> 1);
        return (count & 1u) ? c_mul(EXP_I_PI_4, z) : z;
    }
    case QVNT_RX: case QVNT_RXX: {                             /* rx.rs:18-24, rxx.rs:19-25 */
        orc_c p0 = psi[idx], p1 = psi[idx ^ a], r;
        r.re = p0.re * g->re + p1.im * g->im;
        r.im = p0.im * g->re - p1.re * g->im;
        return r;
    }
    case QVNT_RY: {                                            /* ry.rs:19-29 */
        orc_c ph = {g->re, g->im}, p0 = psi[idx], p1 = psi[idx ^ a], r;
        if ((idx & a) == 0) ph.im = -ph.im;
        r.re = p0.re * ph.re + p1.re * ph.im;
        r.im = p0.im * ph.re + p1.im * ph.im;
        return r;
    }
    case QVNT_RYY: {                                           /* ryy.rs:19-29 */
        orc_c ph = {g->re, g->im}, p0 = psi[idx], p1 = psi[idx ^ a], r;
        if ((popc(idx & a) & 1u) == 0) ph.im = -ph.im;
        r.re = p0.re * ph.re + p1.im * ph.im;
        r.im = p0.im * ph.re - p1.re * ph.im;
        return r;
    }
    case QVNT_RZ: {                                            /* rz.rs:19-25 */
        orc_c ph = {g->re, g->im};
        if ((idx & a) == 0) ph.im = -ph.im;
        return c_mul(ph, psi[idx]);
    }
    case QVNT_RZZ: {                                           /* rzz.rs:19-25 */
        orc_c ph = {g->re, g->im};
        if ((popc(idx & a) & 1u) == 0) ph.im = -ph.im;
        return c_mul(ph, psi[idx]);
    }
    case QVNT_PHI: {                                           /* diag(1, e^{i th}) */
        orc_c ph = {g->re, g->im};
        return (idx & a) ? c_mul(ph, psi[idx]) : psi[idx];
    }
    case QVNT_U1: {                                            /* u1.rs:17-25 */
        uint64_t base = idx & ~a;
        if ((idx & a) == 0)
            return c_add(c_mul(mget(g->m, 0), psi[base]), c_mul(mget(g->m, 1), psi[base | a]));
        return c_add(c_mul(mget(g->m, 2), psi[base]), c_mul(mget(g->m, 3), psi[base | a]));
    }
    case QVNT_U2: {                                            /* u2.rs:22-49 */
        const uint64_t b = g->b;
        int row = ((idx & b) ? 2 : 0) | ((idx & a) ? 1 : 0);
        uint64_t base = idx & ~a & ~b;
        orc_c s = c_mul(mget(g->m, 4 * row + 0), psi[base]);
        s = c_add(s, c_mul(mget(g->m, 4 * row + 1), psi[base | a]));
        s = c_add(s, c_mul(mget(g->m, 4 * row + 2), psi[base | b]));
        s = c_add(s, c_mul(mget(g->m, 4 * row + 3), psi[base | a | b]));
        return s;
    }
    case QVNT_H1: {                                            /* h1.rs:16-22 */
        orc_c p0 = psi[idx], p1 = psi[idx ^ a];
        if (idx & a) p0 = c_neg(p0);
        return c_scale(c_add(p0, p1), FRAC_1_SQRT_2);
    }
    case QVNT_H2: {                                            /* h2.rs:22-38 */
        const uint64_t b = g->b;
        orc_c t0 = psi[idx], t1 = psi[idx ^ a], t2 = psi[idx ^ b], t3 = psi[idx ^ (a | b)];
        if (idx & a) { t0 = c_neg(t0); t2 = c_neg(t2); }
        if (idx & b) { t0 = c_neg(t0); t1 = c_neg(t1); }
        return c_scale(c_add(c_add(c_add(t0, t1), t2), t3), 0.5);
    }
    case QVNT_SWAP:                                            /* swap.rs:16-22 */
        return (popc(idx & a) & 1u) ? psi[idx ^ a] : psi[idx];
    case QVNT_ISWAP: {                                         /* i_swap.rs:20-37 */
        if (popc(idx & a) & 1u) {
            orc_c q = psi[idx ^ a], r;
            if (g->dagger) { r.re = q.im; r.im = -q.re; }
            else { r.re = -q.im; r.im = q.re; }
            return r;
        }
        return psi[idx];
    }
    case QVNT_SQRT_SWAP: {                                     /* sqrt_swap.rs:20-37 */
        if (popc(idx & a) & 1u) {
            orc_c p0 = psi[idx], p1 = psi[idx ^ a], r;
            if (g->dagger) {
                r.re = 0.5 * (p0.re + p0.im + p1.re - p1.im);
                r.im = 0.5 * (p0.im - p0.re + p1.im + p1.re);
            } else {
                r.re = 0.5 * (p0.re - p0.im + p1.re + p1.im);
                r.im = 0.5 * (p0.im + p0.re + p1.im - p1.re);
            }
            return r;
        }
        return psi[idx];
    }
    case QVNT_SQRT_ISWAP: {                                    /* sqrt_i_swap.rs:20-37 */
        if (popc(idx & a) & 1u) {
            orc_c p0 = psi[idx], p1 = psi[idx ^ a], r;
            if (g->dagger) {
                r.re = FRAC_1_SQRT_2 * (p0.re + p1.im);
                r.im = FRAC_1_SQRT_2 * (p0.im - p1.re);
            } else {
                r.re = FRAC_1_SQRT_2 * (p0.re - p1.im);
                r.im = FRAC_1_SQRT_2 * (p0.im + p1.re);
            }
            return r;
        }
        return psi[idx];
    }
    default:
        return psi[idx];
    }
}

/* AtomicOp::for_each / for_each_par (atomic/dispatch.rs:32-67): every
 * output element is gathered; the control predicate `!idx & ctrl == 0`
 * selects the gate, otherwise the input is copied. */
static void for_each(const qvnt_gate* g, const orc_c* psi_i, orc_c* psi_o, uint64_t len, int threads) {
    const uint64_t ctrl = g->ctrl;
    int64_t idx;
    (void)threads;
    if (ctrl != 0) {
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
        for (idx = 0; idx < (int64_t)len; ++idx) {
            uint64_t i = (uint64_t)idx;
            psi_o[i] = ((~i & ctrl) == 0) ? atomic_op(g, psi_i, i) : psi_i[i];
        }
    } else {
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
        for (idx = 0; idx < (int64_t)len; ++idx) psi_o[idx] = atomic_op(g, psi_i, (uint64_t)idx);
    }
}

/* ------------------------------------------------------------ register */
/* QReg::with_state (quant.rs:129-150). */
orc_reg* orc_reg_with_state(uint32_t q_num, uint64_t state) {
    orc_reg* r = (orc_reg*)calloc(1, sizeof(orc_reg));
    uint64_t q_size = (q_num >= 64) ? 0 : (1ull << q_num);
    if (!r) return NULL;
    r->q_num = q_num;
    r->q_mask = q_size - 1u;
    r->len = q_size > MIN_BUFFER_LEN ? q_size : MIN_BUFFER_LEN;
    r->psi = (orc_c*)calloc(r->len, sizeof(orc_c));  /* vec![C_ZERO; len] */
    r->threads = 1;
    if (!r->psi) { free(r); return NULL; }
    r->psi[state & r->q_mask].re = 1.0;
    return r;
}
void orc_reg_free(orc_reg* r) {
    if (!r) return;
    free(r->psi);
    free(r);
}
double* orc_reg_psi(orc_reg* r) { return (double*)r->psi; }
uint64_t orc_reg_len(const orc_reg* r) { return r->len; }

/* QReg::num_threads (quant.rs:186-200): None for 0 or more than the pool. */
int orc_reg_num_threads(orc_reg* r, int threads) {
    int max_threads = 1;
#ifdef _OPENMP
    max_threads = omp_get_num_procs();
#endif
    if (threads <= 0 || threads > max_threads) return -1;
    r->threads = threads;
    return 0;
}

/* QReg::apply (quant.rs:376-395) with MultiOp::apply (multi/mod.rs:96-114):
 * a fresh output buffer, a copy of the input, then each SingleOp gathers
 * psi_i -> psi_o and the two swap. */
void orc_reg_apply(orc_reg* r, const qvnt_gate* g, size_t n) {
    orc_c* psi_o = (orc_c*)malloc(r->len * sizeof(orc_c));   /* set_len, uninit */
    orc_c* psi_i = (orc_c*)malloc(r->len * sizeof(orc_c));   /* psi_i.to_vec() */
    orc_c* t;
    size_t k;
    memcpy(psi_i, r->psi, r->len * sizeof(orc_c));
    for (k = 0; k < n; ++k) {
        for_each(&g[k], psi_i, psi_o, r->len, r->threads);
        t = psi_i; psi_i = psi_o; psi_o = t;
    }
    /* final swap: the result (in psi_i) becomes psi_o, which QReg swaps in. */
    free(r->psi);
    free(psi_o);
    r->psi = psi_i;
}

/* QReg::get_absolute (quant.rs:458-466), Single model: sequential sum of
 * norm_sqr over the whole buffer. */
double orc_reg_get_absolute(const orc_reg* r) {
    double s = 0.0;
    uint64_t i;
    for (i = 0; i < r->len; ++i) s += c_norm_sqr(r->psi[i]);
    return s;
}
/* QReg::get_probabilities (quant.rs:434-454): total over the whole buffer,
 * p_i = norm_sqr_i * (1 / total) for i < 2^n. */
void orc_reg_get_probabilities(const orc_reg* r, double* out) {
    double inv = 1.0 / orc_reg_get_absolute(r);
    uint64_t i, n = 1ull << r->q_num;
    for (i = 0; i < n; ++i) out[i] = c_norm_sqr(r->psi[i]) * inv;
}

/* ------------------------------------------------ measurement contract */
#define LEAF_LOG 6  /* subtrees of 2^6 leaves are recomputed during descent */

static double leaf_w(const orc_c* psi, uint64_t i) {
    double re = psi[i].re, im = psi[i].im;
    return re * re + im * im;
}
/* node(l, j) for l <= LEAF_LOG, computed bottom-up pairwise. */
static double small_node(const orc_c* psi, uint64_t lo, unsigned l) {
    double buf[1u << LEAF_LOG];
    uint64_t m = 1ull << l, k;
    for (k = 0; k < m; ++k) buf[k] = leaf_w(psi, lo + k);
    while (m > 1) {
        for (k = 0; k < m / 2; ++k) buf[k] = buf[2 * k] + buf[2 * k + 1];
        m /= 2;
    }
    return buf[0];
}

typedef struct tree_t {
    unsigned top;      /* log2(count) */
    unsigned base;     /* lowest stored level = min(top, LEAF_LOG) */
    double** lv;       /* lv[l - base][j] = node(l, j) for base <= l <= top */
} tree_t;

static int tree_build(tree_t* t, const orc_c* psi, uint64_t count) {
    unsigned top = 0, l;
    uint64_t j, m;
    while ((1ull << top) < count) ++top;
    t->top = top;
    t->base = top < LEAF_LOG ? top : LEAF_LOG;
    t->lv = (double**)calloc(top - t->base + 1, sizeof(double*));
    m = count >> t->base;
    t->lv[0] = (double*)malloc(m * sizeof(double));
    for (j = 0; j < m; ++j) t->lv[0][j] = small_node(psi, j << t->base, t->base);
    for (l = t->base + 1; l <= top; ++l) {
        double* lo = t->lv[l - 1 - t->base];
        m >>= 1;
        t->lv[l - t->base] = (double*)malloc(m * sizeof(double));
        for (j = 0; j < m; ++j) t->lv[l - t->base][j] = lo[2 * j] + lo[2 * j + 1];
    }
    return 0;
}
static void tree_free(tree_t* t) {
    unsigned l;
    for (l = 0; l <= t->top - t->base; ++l) free(t->lv[l]);
    free(t->lv);
}
static double tree_node(const tree_t* t, const orc_c* psi, unsigned l, uint64_t j) {
    if (l >= t->base) return t->lv[l - t->base][j];
    return small_node(psi, j << l, l);
}

double orc_tree_total(const double* psi_interleaved, uint64_t count) {
    tree_t t;
    double s;
    tree_build(&t, (const orc_c*)psi_interleaved, count);
    s = t.lv[t.top - t.base][0];
    tree_free(&t);
    return s;
}

static uint64_t tree_descend(const tree_t* t, const orc_c* psi, double u, int* err) {
    double total = t->lv[t->top - t->base][0];
    double x;
    uint64_t j = 0;
    unsigned l;
    if (!(total > 0.0) || !isfinite(total)) { *err = QVNT_E_ZERO_NORM; return 0; }
    x = u * total;
    for (l = t->top; l > 0; --l) {
        double L = tree_node(t, psi, l - 1, 2 * j);
        double R = tree_node(t, psi, l - 1, 2 * j + 1);
        if (x < L || R == 0.0) j = 2 * j;
        else { x = x - L; j = 2 * j + 1; }
    }
    *err = 0;
    return j;
}

uint64_t orc_tree_sample(const double* psi_interleaved, uint64_t count, double u, int* err) {
    tree_t t;
    uint64_t r;
    tree_build(&t, (const orc_c*)psi_interleaved, count);
    r = tree_descend(&t, (const orc_c*)psi_interleaved, u, err);
    tree_free(&t);
    return r;
}

/* QReg::measure_mask (quant.rs:490-501) with collapse_mask (quant.rs:468-486).
 * The weights are |psi_i|^2 for i < 2^n (get_probabilities' support); the
 * draw follows the tree contract; collapse runs over the whole buffer. */
uint64_t orc_reg_measure_mask(orc_reg* r, uint64_t mask, double u, uint32_t flags, int* err) {
    uint64_t rand_idx, i;
    *err = 0;
    mask &= r->q_mask;
    if (mask == 0) return 0;                      /* CReg::new(q_num) */
    rand_idx = orc_tree_sample((const double*)r->psi, 1ull << r->q_num, u, err);
    if (*err) return 0;
    for (i = 0; i < r->len; ++i)
        if ((i ^ rand_idx) & mask) { r->psi[i].re = 0.0; r->psi[i].im = 0.0; }
    if (flags & QVNT_MEASURE_RENORM) {
        double kept = orc_tree_total((const double*)r->psi, r->len);
        double scale = 1.0 / sqrt(kept);
        for (i = 0; i < r->len; ++i) r->psi[i] = c_scale(r->psi[i], scale);
    }
    return rand_idx & mask;
}

/* ------------------------------------------------------------------ rng */
static uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
/* SplitMix64 (Steele, Lea, Flood 2014; Vigna's splitmix64.c). */
uint64_t orc_splitmix64(uint64_t* state) {
    *state += 0x9E3779B97F4A7C15ull;
    return mix64(*state);
}
/* key = mix64(seed ^ mix64(stream ^ 0x51564E54)); the draw is output
 * number `index` of SplitMix64 seeded with key, top 53 bits * 2^-53. */
double orc_uniform(uint64_t seed, uint64_t stream, uint64_t index) {
    uint64_t key = mix64(seed ^ mix64(stream ^ 0x51564E54ull));
    uint64_t z = key + (index + 1u) * 0x9E3779B97F4A7C15ull;
    return (double)(mix64(z) >> 11) * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------- register algebra and reset */
/* QReg::reset (quant.rs:202-205): zeros, psi[q_mask & i_state] = 1. */
void orc_reg_reset(orc_reg* r, uint64_t i_state) {
    memset(r->psi, 0, r->len * sizeof(orc_c));
    r->psi[r->q_mask & i_state].re = 1.0;
}

/* QReg::normalize (quant.rs:397-414): norm = sqrt(get_absolute()); reset
 * for norm <= 1e-15; untouched when 1 - norm <= 1e-9 (also for norm > 1);
 * otherwise every amplitude times 1/norm. */
static void normalize(orc_reg* r) {
    double norm = sqrt(orc_reg_get_absolute(r));
    uint64_t i;
    if (norm <= 1e-15) { orc_reg_reset(r, 0); return; }
    if (1. - norm <= 1e-9) return;
    norm = 1. / norm;
    for (i = 0; i < r->len; ++i) r->psi[i] = c_scale(r->psi[i], norm);
}

/* QReg::reset_by_mask (quant.rs:207-229): the whole register -> reset(0);
 * else zero every amplitude with idx & mask != 0, then normalize. */
void orc_reg_reset_by_mask(orc_reg* r, uint64_t mask) {
    uint64_t i;
    if ((mask & r->q_mask) == r->q_mask) { orc_reg_reset(r, 0); return; }
    for (i = 0; i < r->len; ++i)
        if (i & mask) { r->psi[i].re = 0.0; r->psi[i].im = 0.0; }
    normalize(r);
}

static orc_reg* alloc_reg(uint32_t q_num, uint64_t len) {
    orc_reg* r = (orc_reg*)calloc(1, sizeof(orc_reg));
    if (!r) return NULL;
    r->q_num = q_num;
    r->q_mask = (1ull << q_num) - 1u;
    r->len = len;
    r->threads = 1;
    r->psi = (orc_c*)calloc(len, sizeof(orc_c));
    if (!r->psi) { free(r); return NULL; }
    return r;
}

/* tensor_prod (quant.rs:330-371): q_num = a + b; psi[idx] =
 * a[idx & a_mask] * b[(idx >> a.q_num) & b_mask] for idx < 2^q_num, zero
 * in the padding of the max(2^q_num, 8) buffer. */
orc_reg* orc_reg_tensor(const orc_reg* a, const orc_reg* b) {
    uint32_t q_num = a->q_num + b->q_num;
    uint64_t q_size = 1ull << q_num, len = q_size > MIN_BUFFER_LEN ? q_size : MIN_BUFFER_LEN, idx;
    orc_reg* r = alloc_reg(q_num, len);
    if (!r) return NULL;
    for (idx = 0; idx < q_size; ++idx)
        r->psi[idx] = c_mul(a->psi[idx & a->q_mask], b->psi[(idx >> a->q_num) & b->q_mask]);
    return r;
}

/* QReg::combine (quant.rs:245-268): equal sizes; a fills the lower half
 * of a register one qubit larger, b the upper half.  NULL where the
 * reference returns None (sizes differ) or panics (the halves do not fit:
 * fewer than 3 qubits, where both buffers hold the 8-element minimum). */
orc_reg* orc_reg_combine(const orc_reg* a, const orc_reg* b) {
    uint64_t q_size, len;
    orc_reg* r;
    if (a->q_num != b->q_num || a->len != b->len) return NULL;
    q_size = 1ull << (a->q_num + 1);
    len = q_size > MIN_BUFFER_LEN ? q_size : MIN_BUFFER_LEN;
    if (len != 2 * a->len) return NULL;
    r = alloc_reg(a->q_num + 1, len);
    if (!r) return NULL;
    memcpy(r->psi, a->psi, a->len * sizeof(orc_c));
    memcpy(r->psi + a->len, b->psi, b->len * sizeof(orc_c));
    return r;
}

/* QReg::combine_with_unitary (quant.rs:271-308): psi[idx] = c00 a[m & idx]
 * + c01 b[m & idx] where idx has no bit above m (m = a.q_mask), else
 * c10 a[..] + c11 b[..]; c = [c00, c01, c10, c11] interleaved. */
orc_reg* orc_reg_combine_with_unitary(const orc_reg* a, const orc_reg* b, const double c[8]) {
    uint64_t q_size, len, idx, m = a->q_mask;
    orc_reg* r;
    orc_c c00 = {c[0], c[1]}, c01 = {c[2], c[3]}, c10 = {c[4], c[5]}, c11 = {c[6], c[7]};
    if (a->q_num != b->q_num) return NULL;
    q_size = 1ull << (a->q_num + 1);
    len = q_size > MIN_BUFFER_LEN ? q_size : MIN_BUFFER_LEN;
    r = alloc_reg(a->q_num + 1, len);
    if (!r) return NULL;
    for (idx = 0; idx < len; ++idx) {
        orc_c q0 = a->psi[m & idx], q1 = b->psi[m & idx];
        r->psi[idx] = (~m & idx) == 0 ? c_add(c_mul(c00, q0), c_mul(c01, q1)) : c_add(c_mul(c10, q0), c_mul(c11, q1));
    }
    return r;
}

/* QReg::linear_composition (quant.rs:311-328): psi = psi * c0 + other * c1
 * element by element over buffers of equal length; -1 if they differ. */
int orc_reg_linear_composition(orc_reg* r, const orc_reg* other, const double c[4]) {
    orc_c c0 = {c[0], c[1]}, c1 = {c[2], c[3]};
    uint64_t i;
    if (r->len != other->len) return -1;
    for (i = 0; i < r->len; ++i) r->psi[i] = c_add(c_mul(r->psi[i], c0), c_mul(other->psi[i], c1));
    return 0;
}

/* QReg::set_num (quant.rs:156-166): the buffer is resized to exactly 2^q
 * (zero-extended or truncated); the reset branch after it is dead (it
 * compares with the already-updated q_num), so nothing is renormalised. */
int orc_reg_set_num(orc_reg* r, uint32_t q_num) {
    uint64_t q_size = 1ull << q_num;
    orc_c* psi = (orc_c*)calloc(q_size, sizeof(orc_c));
    if (!psi) return -1;
    memcpy(psi, r->psi, (q_size < r->len ? q_size : r->len) * sizeof(orc_c));
    free(r->psi);
    r->psi = psi;
    r->len = q_size;
    r->q_num = q_num;
    r->q_mask = q_size - 1u;
    return 0;
}

/* QReg::get_polar (quant.rs:417-431): num-complex to_polar = (norm, arg) =
 * (hypot(re, im), atan2(im, re)) for the first 2^q amplitudes. */
void orc_reg_get_polar(const orc_reg* r, double* out) {
    uint64_t i, n = 1ull << r->q_num;
    for (i = 0; i < n; ++i) {
        out[2 * i] = hypot(r->psi[i].re, r->psi[i].im);
        out[2 * i + 1] = atan2(r->psi[i].im, r->psi[i].re);
    }
}

/* sample_all (quant.rs:513-594) as an exact multinomial: `shots` seeded
 * draws u_k = orc_uniform(seed, stream, k) through the tree contract over
 * the 2^q weights; hist has one count per buffer element.  The reference's
 * normal approximation with an unseeded generator is not reproducible. */
int orc_reg_sample_all(const orc_reg* r, uint64_t shots, uint64_t seed, uint64_t stream, uint64_t* hist) {
    tree_t t;
    uint64_t k;
    int err = 0;
    memset(hist, 0, r->len * sizeof(uint64_t));
    if (shots == 0) return 0;
    tree_build(&t, r->psi, 1ull << r->q_num);
    for (k = 0; k < shots && !err; ++k) {
        uint64_t j = tree_descend(&t, r->psi, orc_uniform(seed, stream, k), &err);
        if (!err) hist[j] += 1;
    }
    tree_free(&t);
    return err;
}

/* ------------------------------------------------ the QASM driver loop */
/* CReg::get_by_mask (class.rs:106-117): the value's bits at the set bits of
 * mask & q_mask, packed from bit 0 upwards. */
static uint64_t get_by_mask(uint64_t value, uint64_t mask) {
    uint64_t out = 0, m;
    unsigned k = 0;
    for (m = mask; m; m &= m - 1, ++k)
        if (value & (m & (~m + 1))) out |= 1ull << k;
    return out;
}

/* Sym::finish (qasm/sym.rs:40-74) on a register and a classical value of
 * c_num bits: for each step apply its op, then by separator
 *   MEASURE(a = q_arg, b = c_arg): measure_mask(q_arg) with the next draw
 *     (none is taken when q_arg & q_mask == 0); the i-th set bit of c_arg
 *     is set (or xored, measure_xor) from the i-th set bit of q_arg;
 *   IF(a = c, b = v): the op runs only if get_by_mask(c) == v;
 *   RESET(a = q): reset_by_mask(q);
 * and finally the tail op. */
int orc_run_program(orc_reg* r, const orc_step* steps, size_t n, const qvnt_gate* tail, size_t tail_n,
                    uint32_t measure_xor, uint64_t* cval, uint32_t c_num, const double* draws, size_t n_draws,
                    size_t* used) {
    uint64_t c_mask = c_num >= 64 ? ~0ull : ((1ull << c_num) - 1u);
    size_t k;
    *used = 0;
    for (k = 0; k < n; ++k) {
        const orc_step* s = &steps[k];
        if (s->sep == ORC_SEP_IF) {
            if (get_by_mask(*cval, s->a & c_mask) == s->b) orc_reg_apply(r, s->g, s->n);
            continue;
        }
        orc_reg_apply(r, s->g, s->n);
        if (s->sep == ORC_SEP_MEASURE) {
            uint64_t outcome = 0, qm, cm;
            if (s->a & r->q_mask) {
                int err = 0;
                if (*used >= n_draws) return QVNT_E_ARG;
                outcome = orc_reg_measure_mask(r, s->a, draws[(*used)++], 0, &err);
                if (err) return err;
            }
            for (qm = s->a, cm = s->b; qm && cm; qm &= qm - 1, cm &= cm - 1) {
                uint64_t qb = qm & (~qm + 1), cb = cm & (~cm + 1);
                int bit = (outcome & qb) != 0;
                if (measure_xor) { if (bit) *cval ^= cb; }
                else if (bit) *cval |= cb;
                else *cval &= ~cb;
            }
        } else if (s->sep == ORC_SEP_RESET) {
            orc_reg_reset_by_mask(r, s->a);
        }
    }
    orc_reg_apply(r, tail, tail_n);
    return 0;
}
