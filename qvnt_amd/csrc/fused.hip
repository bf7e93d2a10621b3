// The fused form of QReg::apply: many SingleOps per pass over the state.
//
// The host compiler (fusion.cpp) turns the queue into passes; this kernel
// runs one pass.  One block per tile of 2^t amplitudes (the tile's t qubits
// are the pass's qubit set; blocks take tiles in XCD order, xcd_order).
// The tile is read once from HBM, every phase
// moves its 16 amplitudes per thread LDS -> registers -> LDS (skipped when
// the phase's register qubits are the tile's top four, the layout of the
// HBM access itself), and the tile is written back once.  HBM traffic is
// one read and one write of the state per pass, whatever the number of
// gates in it.  Results agree with the plain form to rounding (FMA
// contraction is on in this file; 2x2 products are merged on the host).
#pragma clang fp contract(fast)

#include <hip/hip_cooperative_groups.h>

#include <cstring>
#include <vector>

#include "device.hpp"
#include "fusion.hpp"
#include "reg.hpp"

namespace qvnt {
namespace {

using namespace fusion;

// the LDS swizzle of fusion.hpp (lds_swizzle), for device code
__device__ __forceinline__ uint32_t dswz(uint32_t e) { return e ^ ((e >> 4) & 15u); }

__device__ __forceinline__ double fmar(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmar(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// complex multiply(-add) as explicit FMA chains: 4 instructions each
template <class R>
__device__ __forceinline__ cplx<R> fma_c(cplx<R> acc, cplx<R> a, cplx<R> x) {
    acc.re = fmar(-a.im, x.im, fmar(a.re, x.re, acc.re));
    acc.im = fmar(a.im, x.re, fmar(a.re, x.im, acc.im));
    return acc;
}
template <class R>
__device__ __forceinline__ cplx<R> mul_c(cplx<R> a, cplx<R> x) {
    return {fmar(a.re, x.re, -(a.im * x.im)), fmar(a.re, x.im, a.im * x.re)};
}
template <class R>
__device__ __forceinline__ cplx<R> ld_coef(const double* __restrict__ c, int k) {
    return {(R)c[2 * k], (R)c[2 * k + 1]};
}

// 2x2 on register bit J for the register indices in the uniform pattern
// pc (register controls); MASKED = false when there are none.
template <int J, bool MASKED, class R>
__device__ __forceinline__ void apply_m2(cplx<R>* v, const double* __restrict__ c, uint32_t pc) {
    const cplx<R> u00 = ld_coef<R>(c, 0), u01 = ld_coef<R>(c, 1), u10 = ld_coef<R>(c, 2), u11 = ld_coef<R>(c, 3);
#pragma unroll
    for (int i = 0; i < kRegs; ++i) {
        if (i & (1 << J)) continue;
        if (MASKED && !((pc >> i) & 1u)) continue;
        const cplx<R> x0 = v[i], x1 = v[i | (1 << J)];
        v[i] = fma_c(mul_c(u00, x0), u01, x1);
        v[i | (1 << J)] = fma_c(mul_c(u10, x0), u11, x1);
    }
}

template <int J0, int J1, class R>
__device__ __forceinline__ void apply_m4(cplx<R>* v, const double* __restrict__ c, uint32_t pc) {
#pragma unroll
    for (int i = 0; i < kRegs; ++i) {
        if (i & ((1 << J0) | (1 << J1))) continue;
        if (!((pc >> i) & 1u)) continue;
        const int idx[4] = {i, i | (1 << J0), i | (1 << J1), i | (1 << J0) | (1 << J1)};
        cplx<R> x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = v[idx[k]];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            cplx<R> acc = mul_c(ld_coef<R>(c, 4 * r), x[0]);
            acc = fma_c(acc, ld_coef<R>(c, 4 * r + 1), x[1]);
            acc = fma_c(acc, ld_coef<R>(c, 4 * r + 2), x[2]);
            acc = fma_c(acc, ld_coef<R>(c, 4 * r + 3), x[3]);
            v[idx[r]] = acc;
        }
    }
}

// anti-diagonal 2x2 (u00 = u11 = 0): out0 = u01 x1, out1 = u10 x0
template <int J, bool MASKED, class R>
__device__ __forceinline__ void apply_a2(cplx<R>* v, const double* __restrict__ c, uint32_t pc) {
    const cplx<R> u01 = ld_coef<R>(c, 0), u10 = ld_coef<R>(c, 1);
#pragma unroll
    for (int i = 0; i < kRegs; ++i) {
        if (i & (1 << J)) continue;
        if (MASKED && !((pc >> i) & 1u)) continue;
        const cplx<R> x0 = v[i], x1 = v[i | (1 << J)];
        v[i] = mul_c(u01, x1);
        v[i | (1 << J)] = mul_c(u10, x0);
    }
}

// the swap of the two halves (x, controlled x): no arithmetic
template <int J, bool MASKED, class R>
__device__ __forceinline__ void apply_x2(cplx<R>* v, uint32_t pc) {
#pragma unroll
    for (int i = 0; i < kRegs; ++i) {
        if (i & (1 << J)) continue;
        if (MASKED && !((pc >> i) & 1u)) continue;
        const cplx<R> x0 = v[i];
        v[i] = v[i | (1 << J)];
        v[i | (1 << J)] = x0;
    }
}

// the real middle factor of a split 2x2 (fusion.hpp: OP_T2), one FMA per
// real output: out0 = x0 - a x1, out1 = x1 + a x0
template <int J, bool MASKED, class R>
__device__ __forceinline__ void apply_t2(cplx<R>* v, R a, uint32_t pc) {
#pragma unroll
    for (int i = 0; i < kRegs; ++i) {
        if (i & (1 << J)) continue;
        if (MASKED && !((pc >> i) & 1u)) continue;
        const cplx<R> x0 = v[i], x1 = v[i | (1 << J)];
        v[i] = cplx<R>{fmar(-a, x1.re, x0.re), fmar(-a, x1.im, x0.im)};
        v[i | (1 << J)] = cplx<R>{fmar(a, x0.re, x1.re), fmar(a, x0.im, x1.im)};
    }
}

template <class R>
__device__ __forceinline__ void apply_t2_sel(cplx<R>* v, R a, uint32_t sel, uint32_t pc) {
    if (pc == 0xFFFFu) {
        switch (sel) {
        case 0: apply_t2<0, false>(v, a, 0u); break;
        case 1: apply_t2<1, false>(v, a, 0u); break;
        case 2: apply_t2<2, false>(v, a, 0u); break;
        default: apply_t2<3, false>(v, a, 0u); break;
        }
    } else {
        switch (sel) {
        case 0: apply_t2<0, true>(v, a, pc); break;
        case 1: apply_t2<1, true>(v, a, pc); break;
        case 2: apply_t2<2, true>(v, a, pc); break;
        default: apply_t2<3, true>(v, a, pc); break;
        }
    }
}

// 4x4 pairing r with r ^ 3 only: out[r] = u[r][r] x[r] + u[r][r^3] x[r^3]
template <int J0, int J1, class R>
__device__ __forceinline__ void apply_p4(cplx<R>* v, const double* __restrict__ c, uint32_t pc) {
#pragma unroll
    for (int i = 0; i < kRegs; ++i) {
        if (i & ((1 << J0) | (1 << J1))) continue;
        if (!((pc >> i) & 1u)) continue;
        const int idx[4] = {i, i | (1 << J0), i | (1 << J1), i | (1 << J0) | (1 << J1)};
        cplx<R> x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = v[idx[k]];
#pragma unroll
        for (int r = 0; r < 4; ++r)
            v[idx[r]] = fma_c(mul_c(ld_coef<R>(c, 2 * r), x[r]), ld_coef<R>(c, 2 * r + 1), x[r ^ 3]);
    }
}

// P4 with real diagonal and imaginary pairing coefficients (fusion.hpp:
// OP_P4R): out_r = c_r x_r + i s_r x_(r^3)
template <int J0, int J1, class R>
__device__ __forceinline__ void apply_p4r(cplx<R>* v, const double* __restrict__ c, uint32_t pc) {
#pragma unroll
    for (int i = 0; i < kRegs; ++i) {
        if (i & ((1 << J0) | (1 << J1))) continue;
        if (!((pc >> i) & 1u)) continue;
        const int idx[4] = {i, i | (1 << J0), i | (1 << J1), i | (1 << J0) | (1 << J1)};
        cplx<R> x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = v[idx[k]];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const cplx<R> k = ld_coef<R>(c, r);
            v[idx[r]] = cplx<R>{k.re * x[r].re - k.im * x[r ^ 3].im, k.re * x[r].im + k.im * x[r ^ 3].re};
        }
    }
}

// diagonal: one phase per register (host tables, controls folded in)
template <class R>
__device__ __forceinline__ void apply_d(cplx<R>* v, const double* __restrict__ c) {
#pragma unroll
    for (int i = 0; i < kRegs; ++i) v[i] = mul_c(ld_coef<R>(c, i), v[i]);
}

// a phase on the amplitudes whose register bit J is 1 (fusion.hpp: OP_PH)
template <int J, class R>
__device__ __forceinline__ void apply_ph(cplx<R>* v, const cplx<R> g) {
#pragma unroll
    for (int i = 0; i < kRegs; ++i)
        if (i & (1 << J)) v[i] = mul_c(g, v[i]);
}

// the structured block (fusion.hpp: OP_S): diagonal, T2 on register bits
// 0..3, diagonal, either diagonal left out when the op's sel says so --
// each form is straight-line code, so the 16 amplitudes stay in registers
// from the first product to the last
template <bool D1, bool D2, class R>
__device__ __forceinline__ void apply_s(cplx<R>* v, const double* __restrict__ c) {
    if (D1) apply_d(v, c);
    apply_t2<0, false>(v, (R)c[2 * kRegs + 0], 0u);
    apply_t2<1, false>(v, (R)c[2 * kRegs + 2], 0u);
    apply_t2<2, false>(v, (R)c[2 * kRegs + 4], 0u);
    apply_t2<3, false>(v, (R)c[2 * kRegs + 6], 0u);
    if (D2) apply_d(v, c + 2 * (kRegs + kRegBits));
}


// T2's and PH's coefficient, carried in the op (fusion.hpp: FOp::ga, gb)
template <class R>
__device__ __forceinline__ cplx<R> inline_coef(const FOp& op) {
    return cplx<R>{(R)__longlong_as_double((long long)op.ga), (R)__longlong_as_double((long long)op.gb)};
}

template <class R>
__device__ __forceinline__ void run_phase(cplx<R>* v, const FPhase& ph, const FOp* __restrict__ ops,
                                          const double* __restrict__ coef, uint32_t tb, uint64_t tile) {
    // the product of the phase's G factors (uniform, accumulated in FP64)
    cplx<double> scal{1.0, 0.0};
    bool scaled = false;
    // Each op's header is loaded when the op starts.  Loading the next one
    // ahead held 16 more SGPRs through the op body and doubled the SGPR
    // spills (into VGPR lanes); the whole-circuit passes ran 3% slower.
    for (uint32_t o = ph.op_begin; o < ph.op_end; ++o) {
        const FOp op = ops[o];
        // controls on lane bits (this thread's base) and on bits outside the tile
        if ((tb & op.lctrl) != op.lctrl || (tile & op.gctrl) != op.gctrl) continue;
        const double* c = coef + op.coef;
        if (op.type == OP_G) {
            const uint32_t k = ((tile & op.ga) != 0 ? 1u : 0u) | ((tile & op.gb) != 0 ? 2u : 0u);
            scal = mul_c(ld_coef<double>(c, (int)(k * kRegs)), scal);
            scaled = true;
        } else if (op.type == OP_D) {
            // table variant from the non-register operands (uniform branches)
            const bool ea = (op.la | op.ga) != 0, eb = (op.lb | op.gb) != 0;
            const uint32_t k = ((((tb & op.la) | (tile & op.ga)) != 0) ? 1u : 0u) |
                               ((((tb & op.lb) | (tile & op.gb)) != 0) ? 2u : 0u);
            // dmode 1: the variant-0 table, then (waves whose operand bit is
            // set) the ratio table, both uniform loads; else each lane loads
            // its own variant
            if (!ea && !eb) {
                apply_d(v, c);
            } else if (op.dmode == 1) {
                apply_d(v, c);
                const uint32_t var = ea ? 1u : 2u;
                if (k & var) apply_d(v, c + 2 * kRegs * var);  // uniform branch
            } else {
                apply_d(v, c + 2 * kRegs * k);  // per-lane variant table
            }
        } else if (op.type == OP_S) {
            switch (op.sel) {
            case 0: apply_s<false, false>(v, c); break;
            case 1: apply_s<true, false>(v, c); break;
            case 2: apply_s<false, true>(v, c); break;
            default: apply_s<true, true>(v, c); break;
            }
        } else if (op.type == OP_PH) {
            switch (op.sel) {
            case 0: apply_ph<0>(v, inline_coef<R>(op)); break;
            case 1: apply_ph<1>(v, inline_coef<R>(op)); break;
            case 2: apply_ph<2>(v, inline_coef<R>(op)); break;
            default: apply_ph<3>(v, inline_coef<R>(op)); break;
            }
        } else if (op.type == OP_T2) {
            apply_t2_sel(v, inline_coef<R>(op).re, op.sel, op.pc);
        } else if (op.type == OP_M2) {
            if (op.pc == 0xFFFFu) {
                switch (op.sel) {
                case 0: apply_m2<0, false>(v, c, 0u); break;
                case 1: apply_m2<1, false>(v, c, 0u); break;
                case 2: apply_m2<2, false>(v, c, 0u); break;
                default: apply_m2<3, false>(v, c, 0u); break;
                }
            } else {
                switch (op.sel) {
                case 0: apply_m2<0, true>(v, c, op.pc); break;
                case 1: apply_m2<1, true>(v, c, op.pc); break;
                case 2: apply_m2<2, true>(v, c, op.pc); break;
                default: apply_m2<3, true>(v, c, op.pc); break;
                }
            }
        } else if (op.type == OP_M4) {
            switch (op.sel) {
            case 1: apply_m4<0, 1>(v, c, op.pc); break;
            case 2: apply_m4<0, 2>(v, c, op.pc); break;
            case 3: apply_m4<0, 3>(v, c, op.pc); break;
            case 6: apply_m4<1, 2>(v, c, op.pc); break;
            case 7: apply_m4<1, 3>(v, c, op.pc); break;
            default: apply_m4<2, 3>(v, c, op.pc); break;
            }
        } else if (op.type == OP_P4) {
            switch (op.sel) {
            case 1: apply_p4<0, 1>(v, c, op.pc); break;
            case 2: apply_p4<0, 2>(v, c, op.pc); break;
            case 3: apply_p4<0, 3>(v, c, op.pc); break;
            case 6: apply_p4<1, 2>(v, c, op.pc); break;
            case 7: apply_p4<1, 3>(v, c, op.pc); break;
            default: apply_p4<2, 3>(v, c, op.pc); break;
            }
        } else if (op.type == OP_P4R) {
            switch (op.sel) {
            case 1: apply_p4r<0, 1>(v, c, op.pc); break;
            case 2: apply_p4r<0, 2>(v, c, op.pc); break;
            case 3: apply_p4r<0, 3>(v, c, op.pc); break;
            case 6: apply_p4r<1, 2>(v, c, op.pc); break;
            case 7: apply_p4r<1, 3>(v, c, op.pc); break;
            default: apply_p4r<2, 3>(v, c, op.pc); break;
            }
        } else if (op.type == OP_A2) {
            if (op.pc == 0xFFFFu) {
                switch (op.sel) {
                case 0: apply_a2<0, false>(v, c, 0u); break;
                case 1: apply_a2<1, false>(v, c, 0u); break;
                case 2: apply_a2<2, false>(v, c, 0u); break;
                default: apply_a2<3, false>(v, c, 0u); break;
                }
            } else {
                switch (op.sel) {
                case 0: apply_a2<0, true>(v, c, op.pc); break;
                case 1: apply_a2<1, true>(v, c, op.pc); break;
                case 2: apply_a2<2, true>(v, c, op.pc); break;
                default: apply_a2<3, true>(v, c, op.pc); break;
                }
            }
        } else {
            if (op.pc == 0xFFFFu) {
                switch (op.sel) {
                case 0: apply_x2<0, false>(v, 0u); break;
                case 1: apply_x2<1, false>(v, 0u); break;
                case 2: apply_x2<2, false>(v, 0u); break;
                default: apply_x2<3, false>(v, 0u); break;
                }
            } else {
                switch (op.sel) {
                case 0: apply_x2<0, true>(v, op.pc); break;
                case 1: apply_x2<1, true>(v, op.pc); break;
                case 2: apply_x2<2, true>(v, op.pc); break;
                default: apply_x2<3, true>(v, op.pc); break;
                }
            }
        }
    }
    if (scaled) {
        const cplx<R> f{(R)scal.re, (R)scal.im};
#pragma unroll
        for (int i = 0; i < kRegs; ++i) v[i] = mul_c(f, v[i]);
    }
}

typedef double mf64x4 __attribute__((ext_vector_type(4)));
typedef float mf32x4 __attribute__((ext_vector_type(4)));

// One 16x16x4 matrix-core step: acc(16x16) += A(16x4) * B(4x16).
// Lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16].  Element i of
// its accumulator is acc[l / 16 + 4 i][l % 16] for f64 and
// acc[4 (l / 16) + i][l % 16] for f32 (measured on gfx950).
__device__ __forceinline__ mf64x4 mfma(double a, double b, mf64x4 acc) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
}
__device__ __forceinline__ mf32x4 mfma(float a, float b, mf32x4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
}
template <class R> struct AccOf;
template <> struct AccOf<double> { typedef mf64x4 type; static constexpr bool hi_low_bits = true; };
template <> struct AccOf<float> { typedef mf32x4 type; static constexpr bool hi_low_bits = false; };

// swizzled LDS offset of register-index bits (j & 3) placed at register
// bits (b0, b0 + 1) of the phase: the swizzle is XOR-linear, so offsets of
// disjoint bit groups combine by XOR
__device__ __forceinline__ uint32_t reg_off2(const FPhase& ph, int b0, uint32_t j) {
    return dswz(((j & 1u) << ph.rpos[b0]) | (((j >> 1) & 1u) << ph.rpos[b0 + 1]));
}

// A dense phase: the tile (in LDS) is a 16 x 2^(t-4) complex matrix X
// (row = register index of the phase, column = thread base), and the phase
// computes X := U X as a real 32x32 GEMM on the matrix cores:
//   Yre = Ure Xre - Uim Xim,   Yim = Ure Xim + Uim Xre.
// Each 64-lane wave owns the 64 columns of its own threads (as in the
// register phases), in four 16-column tiles; all reads of a tile's columns
// precede its writes, so the update is in place.
template <class R>
__device__ __forceinline__ void dense_phase(cplx<R>* lds, const FPhase& ph, const double* __restrict__ u, uint32_t tid) {
    typedef typename AccOf<R>::type Acc;
    const uint32_t lane = tid & 63u, hi = lane >> 4, wave0 = tid & ~63u;
    // A operands: U[m][4s + hi] with m = lane % 16
    R ure[4], uim[4], nuim[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int e = 16 * (int)(lane & 15u) + 4 * s + (int)hi;
        ure[s] = (R)u[2 * e];
        uim[s] = (R)u[2 * e + 1];
        nuim[s] = -uim[s];
    }
    const uint32_t in_hi = reg_off2(ph, 0, hi);   // B rows 4s + hi: register bits 0,1 = hi
    // accumulator rows: hi + 4i (f64) or 4 hi + i (f32)
    constexpr bool lowhi = AccOf<R>::hi_low_bits;
    const uint32_t out_hi = reg_off2(ph, lowhi ? 0 : 2, hi);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        // this lane's column and its LDS base (zeros at the register bits)
        uint32_t tb = wave0 + 16u * c + (lane & 15u);
#pragma unroll
        for (int k = 0; k < kRegBits; ++k) {
            const uint32_t lo = tb & ((1u << ph.rpos[k]) - 1u);
            tb = ((tb ^ lo) << 1) | lo;
        }
        const uint32_t stb = dswz(tb);
        cplx<R> x[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) x[s] = lds[stb ^ in_hi ^ reg_off2(ph, 2, (uint32_t)s)];
        Acc yre = {0, 0, 0, 0}, yim = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            yre = mfma(ure[s], x[s].re, yre);
            yim = mfma(ure[s], x[s].im, yim);
            yre = mfma(nuim[s], x[s].im, yre);
            yim = mfma(uim[s], x[s].re, yim);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) lds[stb ^ out_hi ^ reg_off2(ph, lowhi ? 2 : 0, (uint32_t)i)] = cplx<R>{yre[i], yim[i]};
    }
}

// Global index of local element 0 of tile `tile_index` (the tiles
// enumerate the bits outside the tile's qubit set): zeros inserted at the
// tile's bit positions.
__device__ __forceinline__ uint64_t tile_base(const uint32_t* tp, uint32_t t, uint64_t tile_index) {
    uint64_t tile = tile_index;
#pragma unroll
    for (uint32_t k = 0; k < (uint32_t)kMaxTile; ++k) {
        if (k < t) {  // uniform
            const uint64_t lo = tile & ((1ull << tp[k]) - 1u);
            tile = ((tile ^ lo) << 1) | lo;
        }
    }
    return tile;
}

// The pass's local-bit positions from the packed copy (one byte each):
// four words read at once instead of a dependent scalar load per bit in
// the loops over them, and four scalar registers instead of thirteen.
__device__ __forceinline__ void load_tpos(const FPass& pass, uint32_t* tp) {
    uint32_t w[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] = pass.tpack[i];
#pragma unroll
    for (int k = 0; k < kMaxTile; ++k) tp[k] = (w[k / 4] >> (8 * (k % 4))) & 63u;
}

// This thread's swizzled LDS offset in the direct layout, made opaque at
// its use: the 16 offsets derived from it are used at most twice per tile,
// and keeping them live across the phases (where the compiler would hoist
// them) costs 16 VGPRs for the whole kernel.
__device__ __forceinline__ uint32_t direct_base(uint32_t tid) {
    uint32_t s = dswz(tid);
    asm volatile("" : "+v"(s));
    return s;
}

// The tile's final read-out from LDS into the direct layout (register j =
// local tid + j * nth), through the pass's qubit permutation if it has one.
// KARG: the pass is a kernel argument (k_fused): its tables are read bit by
// bit, in a loop -- fully unrolled, the compiler hoists those loads to the
// kernel's start and the extra live scalar registers spill.  Passes in
// global memory (the batched kernels) take the unrolled form: one wide
// scalar load instead of a dependent load per bit.
template <class R, bool KARG>
__device__ __forceinline__ void read_out(cplx<R>* v, const cplx<R>* lds, const FPass& pass, uint32_t tid) {
    if (pass.permuted) {
        uint32_t s = 0;
        if (KARG) {
            for (uint32_t k = 0; k + kRegBits < pass.t; ++k)
                if (tid >> k & 1) s ^= pass.pimg[k];
        } else {
            // (tid's bits from t-4 up and pimg's entries there are zero)
#pragma unroll
            for (uint32_t k = 0; k < (uint32_t)(kMaxTile - kRegBits); ++k) s ^= pass.pimg[k] & (0u - ((tid >> k) & 1u));
        }
#pragma unroll
        for (int j = 0; j < kRegs; ++j) v[j] = lds[s ^ pass.pout[j]];
    } else {
        const uint32_t nth = pass.nthreads;
#pragma unroll
        for (int j = 0, sdir = direct_base(tid); j < kRegs; ++j) v[j] = lds[sdir ^ dswz(j * nth)];
    }
}

// The phases of one pass on one tile.  On entry and on exit v holds the
// thread's amplitudes in the direct layout (register j = local tid + j *
// nth); `lds` is the tile's swizzled LDS copy (local element e at dswz(e)).
template <class R, bool KARG>
__device__ __forceinline__ void tile_phases(cplx<R>* v, cplx<R>* lds, const FPass& pass, const FPhase* __restrict__ phases,
                                            const FOp* __restrict__ ops, const double* __restrict__ coef, uint64_t tile,
                                            uint32_t tid) {
    const uint32_t nth = pass.nthreads;
    const uint32_t p0 = pass.phase_begin, p1 = pass.phase_end;
    for (uint32_t p = p0; p < p1; ++p) {
        const FPhase& ph = phases[p];
        if (ph.dense) {
            // works on the LDS tile: bring the HBM registers in first, and
            // take the result back out for the final store
            if (p == p0) {
#pragma unroll
                for (int j = 0, sdir = direct_base(tid); j < kRegs; ++j) lds[sdir ^ dswz(j * nth)] = v[j];
                __syncthreads();
            }
            dense_phase<R>(lds, ph, coef + ph.ucoef, tid);
            __syncthreads();
            if (p + 1 == p1) read_out<R, KARG>(v, lds, pass, tid);
            continue;
        }
        // local base of this thread in the phase: zeros at the register bits
        uint32_t tb = tid;
#pragma unroll
        for (int k = 0; k < kRegBits; ++k) {
            const uint32_t lo = tb & ((1u << ph.rpos[k]) - 1u);
            tb = ((tb ^ lo) << 1) | lo;
        }
        const uint32_t stb = dswz(tb);
        const bool direct_in = (p == p0) && pass.direct_first;
        const bool direct_out = (p + 1 == p1) && pass.direct_last;
        if (!direct_in) {
            if (p == p0) {
#pragma unroll
                for (int j = 0, sdir = direct_base(tid); j < kRegs; ++j) lds[sdir ^ dswz(j * nth)] = v[j];
                __syncthreads();
            }
#pragma unroll
            for (int i = 0; i < kRegs; ++i) v[i] = lds[stb ^ ph.sroff[i]];
        }
        run_phase(v, ph, ops, coef, tb, tile);
        if (!direct_out) {
#pragma unroll
            for (int i = 0; i < kRegs; ++i) lds[stb ^ ph.sroff[i]] = v[i];
            __syncthreads();
            if (p + 1 == p1) read_out<R, KARG>(v, lds, pass, tid);
        }
    }
}

// global offset of this thread's local bits 0..t-5
__device__ __forceinline__ uint64_t thread_offset(const uint32_t* tp, uint32_t tid) {
    uint64_t lo_off = 0;
    // tid < 2^(t-4): its bits from t-4 up are zero
#pragma unroll
    for (uint32_t k = 0; k < (uint32_t)(kMaxTile - kRegBits); ++k) lo_off |= (uint64_t)((tid >> k) & 1u) << tp[k];
    return lo_off;
}
// every amplitude of the tile is read once and written once per pass:
// non-temporal stores (streamed through the caches, evicted first)
template <class R>
__device__ __forceinline__ void store_direct(cplx<R>* __restrict__ psi, const FPass& pass, uint64_t base, const cplx<R>* v) {
#pragma unroll
    for (int j = 0; j < kRegs; ++j) {
        cplx<R>* p = psi + base + pass.hoff[j];
        __builtin_nontemporal_store(v[j].re, &p->re);
        __builtin_nontemporal_store(v[j].im, &p->im);
    }
}

// One tile of one pass: read 2^t amplitudes, run the pass's phases, write
// them back; the pass and program live in scalar-loadable memory.
// tid: the thread's index within the tile (0 .. 2^(t-4) - 1).  ACTIVE
// false: the threads run the phases (and their barriers) on their own LDS
// region but neither read nor write HBM -- a block of the persistent
// batched kernel whose tile group has no tile (see k_batch_persist).
template <class R, bool KARG, bool GATED = false>
__device__ __forceinline__ void fused_tile(cplx<R>* __restrict__ psi, const FPass& pass, const FPhase* __restrict__ phases,
                                           const FOp* __restrict__ ops, const double* __restrict__ coef,
                                           uint64_t tile_index, cplx<R>* lds, uint32_t tid, bool active = true) {
    uint32_t tp[kMaxTile];
    load_tpos(pass, tp);
    const uint64_t tile = tile_base(tp, pass.t, tile_index);
    const uint64_t base = tile | thread_offset(tp, tid);
    cplx<R> v[kRegs];
#pragma unroll
    for (int j = 0; j < kRegs; ++j) {
        v[j] = cplx<R>{(R)0, (R)0};
        if (!GATED || active) {
            const cplx<R>* p = psi + base + pass.hoff[j];
            v[j] = cplx<R>{__builtin_nontemporal_load(&p->re), __builtin_nontemporal_load(&p->im)};
        }
    }
    tile_phases<R, KARG>(v, lds, pass, phases, ops, coef, tile, tid);
    if (!GATED || active) store_direct(psi, pass, base, v);
}

// At most two waves per SIMD: the 64 KiB LDS tile allows no more, and the
// whole VGPR file then holds the 16 amplitudes and the op operands
// without spills.  (QVNT_FUSED_WAVES, a build-time knob, asks for that many
// waves per SIMD instead: the VGPR budget shrinks to 512 / waves.)
#ifdef QVNT_FUSED_WAVES
#define QVNT_FUSED_BOUNDS __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(QVNT_FUSED_WAVES, QVNT_FUSED_WAVES)))
#else
#define QVNT_FUSED_BOUNDS __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(1, 2)))
#endif

// XCD-aware block order.  The dispatcher hands workgroups to the eight XCDs
// round robin (block b to XCD b mod 8), and each XCD has its own L2 and
// address translation.  With G = 8q + r blocks, block b < 8q runs work
// item (b mod 8) q + b / 8, so XCD x walks the contiguous range
// [x q, (x + 1) q): its concurrently running blocks touch neighbouring
// lines and pages.  The last r blocks keep their index.  A bijection on
// [0, G).
__device__ __forceinline__ uint32_t xcd_order(uint32_t b, uint32_t G) {
    const uint32_t q = G >> 3;
    return b < (q << 3) ? (b & 7u) * q + (b >> 3) : b;
}

template <class R>
__global__ QVNT_FUSED_BOUNDS void k_fused(cplx<R>* __restrict__ psi, FPass pass, const FPhase* __restrict__ phases,
                                               const FOp* __restrict__ ops, const double* __restrict__ coef) {
    extern __shared__ unsigned char smem[];
    fused_tile<R, true>(psi, pass, phases, ops, coef, xcd_order(blockIdx.x, gridDim.x), reinterpret_cast<cplx<R>*>(smem),
                        threadIdx.x);
}

// The batched form: the tiles of one pass of many registers in one launch
// (every item's pass has the launch's t).  Block b (in XCD order,
// xcd_order) runs batch_tiles tiles of item map[b], from tile
// (b - items[map[b]].block0) * batch_tiles, one after the other; registers,
// passes and programs are addressed by offsets from the two base pointers.
template <class R>
__global__ QVNT_FUSED_BOUNDS void k_batch(cplx<R>* __restrict__ arena, const unsigned char* __restrict__ prog,
                                               const BItem* __restrict__ items, const uint32_t* __restrict__ map) {
    extern __shared__ unsigned char smem[];
    const uint32_t b = xcd_order(blockIdx.x, gridDim.x);
    const BItem it = items[map[b]];
    const FPass& pass = *reinterpret_cast<const FPass*>(prog + it.pass_off);
    constexpr uint32_t per_block = batch_tiles(sizeof(R) == 8);
    const uint32_t first = (b - it.block0) * per_block;
    for (uint32_t k = 0; k < per_block; ++k) {
        const uint32_t tile = first + k;
        if (tile >= it.tiles) break;  // uniform
        // the tile's first LDS writes may hit elements other threads still
        // read in the previous tile's last phase
        if (k) __syncthreads();
        fused_tile<R, false>(arena + it.psi_off, pass, reinterpret_cast<const FPhase*>(prog + it.phase_off),
                             reinterpret_cast<const FOp*>(prog + it.op_off),
                             reinterpret_cast<const double*>(prog + it.coef_off), tile, reinterpret_cast<cplx<R>*>(smem),
                             threadIdx.x);
    }
}

// One work unit of the persistent batched kernel (below): up to
// 2^(tmax-t) tiles of one register pass of tile size t, one tile per group
// of 2^(t-4) threads with its own 2^t-amplitude LDS region; all groups run
// the same program, so they pass the same barriers, and a group without a
// tile runs the phases without HBM traffic.  Groups of whole waves (t >= 10)
// take their tile index wave-uniform, as in k_batch, so the program's
// tile-dependent branches and table loads stay scalar.
template <class R>
__device__ __forceinline__ void persist_unit(cplx<R>* __restrict__ arena, const unsigned char* __restrict__ prog,
                                             const BItem* __restrict__ items, const PUnit& un, unsigned char* smem) {
    const BItem it = items[un.item];
    const FPass& pass = *reinterpret_cast<const FPass*>(prog + it.pass_off);
    const FPhase* ph = reinterpret_cast<const FPhase*>(prog + it.phase_off);
    const FOp* ops = reinterpret_cast<const FOp*>(prog + it.op_off);
    const double* coef = reinterpret_cast<const double*>(prog + it.coef_off);
    const uint32_t tbits = pass.t - kRegBits;                 // threads per tile: 2^tbits
    const uint32_t tid = threadIdx.x & ((1u << tbits) - 1u);
    if (tbits >= 6) {
        const uint32_t group = __builtin_amdgcn_readfirstlane(threadIdx.x >> tbits);
        const uint32_t tile = un.tile0 + group;
        fused_tile<R, false, true>(arena + it.psi_off, pass, ph, ops, coef, tile,
                                   reinterpret_cast<cplx<R>*>(smem) + ((size_t)group << pass.t), tid, tile < un.tile_end);
    } else {
        const uint32_t group = threadIdx.x >> tbits;
        const uint32_t tile = un.tile0 + group;
        fused_tile<R, false, true>(arena + it.psi_off, pass, ph, ops, coef, tile,
                                   reinterpret_cast<cplx<R>*>(smem) + ((size_t)group << pass.t), tid, tile < un.tile_end);
    }
}

// Units of each round (in thousandths) that blocks claim instead of taking
// them in a fixed order, kPersistChunk consecutive units per claim
// (4096-register batch, FP64: 54.2 ms per run with none; one unit per
// claim: 51.6 with 100, 50.7 with 150, 53.1 with 250; with 150, two units
// per claim 48.3, four 48.3; 300 in fours 48.3, 400 in fours 48.5).
constexpr uint32_t kPersistClaimed = 150;
constexpr uint32_t kPersistChunk = 4;

// The batched form as one persistent launch (cooperative: every block is
// resident).  The blocks walk the rounds in order and meet at a grid
// barrier between rounds, which also makes a round's HBM writes visible to
// the next (cooperative groups' grid sync: release and acquire fences at
// device scope).  In each round, a whole number of strides of the round's
// first units go out in a fixed order: with a grid of 8q blocks, XCD x (the
// blocks b with b mod 8 = x, which the dispatcher places on XCD x) walks
// its own contiguous eighth of them, its q blocks q apart.  The rest (about
// kPersistClaimed thousandths) are claimed kPersistChunk consecutive units
// at a time from the round's counter, so blocks that finish early take
// more of the round's tail.
// status[1 + r] is round r's counter: zero when the launch starts, zeroed
// for round r + 1 by block 0 during round r (nobody claims from it before
// the barrier that ends round r), and round 0's zeroed again by block 0 at
// the end, after every block has left it; so the counters are zero again
// when the launch ends.
//
// The dynamic LDS array starts 16-byte aligned after the claim word: the
// tile's 16-byte accesses need it (an array at byte 4 made every run 35 %
// slower).
//
// The launch bounds are k_fused's: they leave the register allocation at
// what the kernel uses, which is what the runtime's occupancy query (and so
// the cooperative grid size) counts.
//
// A launch that is not cooperative (a graph node whose cooperative
// attribute the runtime did not honour) has no grid barrier: every block
// sees that alike, block 0 sets *status to 1, and all return before any
// barrier or HBM access.
template <class R>
__global__ QVNT_FUSED_BOUNDS void k_batch_persist(cplx<R>* __restrict__ arena, const unsigned char* __restrict__ prog,
                                                  const BItem* __restrict__ items, const PUnit* __restrict__ units,
                                                  const uint32_t* __restrict__ round_begin, uint32_t rounds,
                                                  uint32_t* status) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t claim;
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    if (!grid.is_valid()) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *status = 1u;
        return;
    }
    uint32_t* count = status + 1;
    const uint32_t G = gridDim.x;
    const bool by_xcd = (G & 7u) == 0;
    const uint32_t xcd = blockIdx.x & 7u;
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t rb = round_begin[r], u_end = round_begin[r + 1], n = u_end - rb;
        if (blockIdx.x == 0 && threadIdx.x == 0 && r + 1 < rounds) count[r + 1] = 0u;
        const uint32_t claimed = (uint32_t)(((uint64_t)n * kPersistClaimed) / 1000u);
        const uint32_t fixed = ((n - claimed) / G) * G;       // whole strides
        // this block's units of the fixed part: [next, stop) in steps of step
        uint32_t next = rb + blockIdx.x, stop = rb + fixed, step = G;
        if (by_xcd) {
            next = rb + xcd * (fixed >> 3) + (blockIdx.x >> 3);
            stop = rb + (xcd + 1u) * (fixed >> 3);
            step = G >> 3;
        }
        // the claimed units in hand: [have, have_end)
        uint32_t have = 0, have_end = 0;
        for (;;) {
            uint32_t u;
            if (next < stop) {  // uniform
                u = next;
                next += step;
            } else {
                if (have >= have_end) {
                    if (threadIdx.x == 0) claim = atomicAdd(&count[r], 1u);
                    __syncthreads();
                    const uint64_t first = (uint64_t)claim * kPersistChunk;
                    __syncthreads();  // before the next claim overwrites it
                    if (first >= n - fixed) break;
                    have = rb + fixed + (uint32_t)first;
                    have_end = rb + fixed + (uint32_t)min((uint64_t)(n - fixed), first + kPersistChunk);
                }
                u = have++;
            }
            persist_unit<R>(arena, prog, items, units[u], smem);
            // the next unit's first LDS writes may hit elements other
            // threads still read in this unit's last phase
            __syncthreads();
        }
        grid.sync();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) count[0] = 0u;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct Layout {
    size_t ph_off, op_off, cf_off, total;
};
Layout layout_of(const Program& prog) {
    Layout L;
    L.ph_off = 0;
    L.op_off = align256(L.ph_off + prog.phases.size() * sizeof(FPhase));
    L.cf_off = align256(L.op_off + prog.ops.size() * sizeof(FOp));
    L.total = L.cf_off + prog.coef.size() * sizeof(double) + 256;
    return L;
}

}  // namespace

size_t program_bytes(const Program& prog) { return layout_of(prog).total; }

int batch_persist_grid(qvnt_prec prec, uint32_t tmax, uint32_t* blocks) {
    const size_t amp = prec == QVNT_F64 ? 16 : 8;
    const int threads = 1 << (tmax - kRegBits);
    const size_t lds = ((size_t)1 << tmax) * amp;
    int per_cu = 0, dev = 0, cus = 0, coop = 0;
    QVNT_HIP(hipGetDevice(&dev));
    QVNT_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    if (!coop) return fail(QVNT_E_HIP, "the device does not support cooperative launches");
    QVNT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (prec == QVNT_F64)
        QVNT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_persist<double>, threads, lds));
    else
        QVNT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_persist<float>, threads, lds));
    if (per_cu <= 0) return fail(QVNT_E_HIP, "the persistent batched kernel does not fit a compute unit");
    *blocks = (uint32_t)per_cu * (uint32_t)cus;
    return QVNT_OK;
}

int launch_batch_persist(hipStream_t stream, qvnt_prec prec, void* arena, const void* prog, const BItem* items,
                         const PUnit* units, const uint32_t* round_begin, uint32_t rounds, uint32_t tmax,
                         uint32_t blocks, uint32_t* status, hipGraph_t* graph) {
    const size_t amp = prec == QVNT_F64 ? 16 : 8;
    const dim3 grid(blocks), block(1u << (tmax - kRegBits));
    const size_t lds = ((size_t)1 << tmax) * amp;
    const unsigned char* p = static_cast<const unsigned char*>(prog);
    void* args[] = {&arena, &p, &items, &units, &round_begin, &rounds, &status};
    void* fn = prec == QVNT_F64 ? reinterpret_cast<void*>(k_batch_persist<double>)
                                : reinterpret_cast<void*>(k_batch_persist<float>);
    if (!graph) {
        QVNT_HIP(hipLaunchCooperativeKernel(fn, grid, block, args, (unsigned)lds, stream));
        return QVNT_OK;
    }
    // the cooperative launch captured from the stream into a graph (on
    // ROCm 7.2 a captured cooperative launch stays cooperative, while a
    // kernel node built by hand with the cooperative attribute does not;
    // the kernel's own check covers either)
    QVNT_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    const hipError_t e = hipLaunchCooperativeKernel(fn, grid, block, args, (unsigned)lds, stream);
    hipGraph_t g = nullptr;
    const hipError_t end = hipStreamEndCapture(stream, &g);
    if (e != hipSuccess || end != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return fail(QVNT_E_HIP, "capturing the persistent batched launch");
    }
    *graph = g;
    return QVNT_OK;
}

int launch_batch(hipStream_t stream, qvnt_prec prec, void* arena, const void* prog, const BItem* items, const uint32_t* map,
                 uint32_t t, uint32_t blocks) {
    const size_t amp = prec == QVNT_F64 ? 16 : 8;
    const dim3 grid(blocks), block(1u << (t - kRegBits));
    const size_t lds = ((size_t)1 << t) * amp;
    const unsigned char* p = static_cast<const unsigned char*>(prog);
    if (prec == QVNT_F64)
        hipLaunchKernelGGL(k_batch<double>, grid, block, lds, stream, static_cast<c64*>(arena), p, items, map);
    else
        hipLaunchKernelGGL(k_batch<float>, grid, block, lds, stream, static_cast<c32*>(arena), p, items, map);
    QVNT_HIP(hipGetLastError());
    return QVNT_OK;
}

void pack_program(const Program& prog, unsigned char* host) {
    const Layout L = layout_of(prog);
    std::memset(host, 0, L.total);
    if (!prog.phases.empty()) std::memcpy(host + L.ph_off, prog.phases.data(), prog.phases.size() * sizeof(FPhase));
    if (!prog.ops.empty()) std::memcpy(host + L.op_off, prog.ops.data(), prog.ops.size() * sizeof(FOp));
    if (!prog.coef.empty()) std::memcpy(host + L.cf_off, prog.coef.data(), prog.coef.size() * sizeof(double));
}

int launch_program(qvnt_reg* r, const Program& prog, const void* dev) {
    const Layout L = layout_of(prog);
    const unsigned char* d = static_cast<const unsigned char*>(dev);
    const FPhase* dph = reinterpret_cast<const FPhase*>(d + L.ph_off);
    const FOp* dop = reinterpret_cast<const FOp*>(d + L.op_off);
    const double* dcf = reinterpret_cast<const double*>(d + L.cf_off);
    for (const FPass& fp : prog.passes) {
        const uint64_t tiles = r->len >> fp.t;
        const size_t lds = ((size_t)1 << fp.t) * r->amp_bytes();
        Timed t(r, QVNT_K_FUSED, 2.0 * (double)r->amp_bytes() * (double)r->len);
        if (r->prec == QVNT_F64)
            hipLaunchKernelGGL(k_fused<double>, dim3((unsigned)tiles), dim3(fp.nthreads), lds, r->stream,
                               static_cast<c64*>(r->psi), fp, dph, dop, dcf);
        else
            hipLaunchKernelGGL(k_fused<float>, dim3((unsigned)tiles), dim3(fp.nthreads), lds, r->stream,
                               static_cast<c32*>(r->psi), fp, dph, dop, dcf);
        QVNT_HIP(hipGetLastError());
    }
    return QVNT_OK;
}

namespace {

// Uploads the program (stream-ordered, from pageable memory) and launches
// one kernel per pass.
int run_program(qvnt_reg* r, const Program& prog) {
    static thread_local std::vector<unsigned char> host;
    const size_t total = program_bytes(prog);
    if (total > r->prog_cap) {
        QVNT_HIP(hipStreamSynchronize(r->stream));
        void* old = r->prog;
        const size_t old_cap = r->prog_cap;
        r->prog = nullptr;
        r->prog_cap = 0;
        // (a register that owns its memory takes its blocks from the device's
        // cache and gives them back there; the stream is synchronized above)
        if (r->owns_memory) {
            dev_release(r->device, old, old_cap);
            if (dev_alloc(r->device, &r->prog, total) != QVNT_OK)
                return fail(QVNT_E_OOM, "device memory for the fused program");
        } else {
            if (old) QVNT_HIP(hipFree(old));
            if (hipMalloc(&r->prog, total) != hipSuccess) return fail(QVNT_E_OOM, "device memory for the fused program");
        }
        r->prog_cap = total;
    }
    host.resize(total);
    pack_program(prog, host.data());
    QVNT_HIP(hipMemcpyAsync(r->prog, host.data(), total, hipMemcpyHostToDevice, r->stream));
    return launch_program(r, prog, r->prog);
}

}  // namespace

// The steps of a fused apply on 2^ns (>= 2^kMinFused) amplitudes; returns
// the tile size.  max_qubits = 0 and no dense phases from kAutoTileQubits:
// the tile the run-time model picks; otherwise max_qubits (0: the default)
// clamped to what the precision's LDS tile and the buffer allow.
unsigned compile_fused(const qvnt_gate* g, size_t n, unsigned ns, bool fp64, uint32_t max_qubits, uint32_t dense_min,
                       std::vector<Step>& steps) {
    const bool single_plain = ns >= kSinglePlainQubits;
    if (!max_qubits && dense_min == 0 && ns >= kAutoTileQubits)
        return plan_auto(g, n, ns, fp64, steps, single_plain, true);
    const unsigned cap = max_tile(fp64);
    unsigned tq = max_qubits ? max_qubits : kDefaultTile;
    if (tq > cap) tq = cap;
    if (tq < (unsigned)kMinFused) tq = kMinFused;
    if (tq > ns) tq = ns;
    plan(g, n, ns, tq, steps, dense_min, single_plain, true, fp64 ? 16 : 8);
    return tq;
}

void fused_plan_stats(const qvnt_gate* g, size_t n, unsigned ns, bool fp64, uint32_t max_qubits, uint32_t dense_min,
                      double* out) {
    out[0] = out[1] = out[2] = out[3] = 0.0;
    if (ns < (unsigned)kMinFused) {  // the plain kernels, one launch per gate
        out[0] = (double)n;
        return;
    }
    std::vector<Step> steps;
    out[1] = compile_fused(g, n, ns, fp64, max_qubits, dense_min, steps);
    for (const Step& st : steps) {
        if (st.plain) {
            out[0] += 1.0;
            continue;
        }
        out[0] += (double)st.prog.passes.size();
        out[2] += (double)program_cost(st.prog);
        for (const FPhase& ph : st.prog.phases) out[3] += ph.dense;
    }
}

int apply_fused(qvnt_reg* r, const qvnt_gate* g, size_t n, uint32_t max_qubits, uint32_t* passes) {
    if (passes) *passes = 0;
    if (r->ns < (unsigned)kMinFused) {
        for (size_t i = 0; i < n; ++i)
            if (int rc = launch_gate(r, g[i])) return rc;
        if (passes) *passes = (uint32_t)n;
        return QVNT_OK;
    }
    std::vector<Step> steps;
    compile_fused(g, n, r->ns, r->prec == QVNT_F64, max_qubits, r->dense_min, steps);
    uint32_t count = 0;
    for (const Step& st : steps) {
        if (st.plain) {
            if (int rc = launch_gate(r, st.gate)) return rc;
            ++count;
        } else {
            if (int rc = run_program(r, st.prog)) return rc;
            count += (uint32_t)st.prog.passes.size();
        }
    }
    if (passes) *passes = count;
    return QVNT_OK;
}

}  // namespace qvnt
