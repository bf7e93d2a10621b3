// A CPU model of the fused kernel (qvnt_amd/csrc/fused.hip: fused_tile,
// run_phase, dense_phase) driven by the real host compiler
// (qvnt_amd/csrc/fusion.cpp: fusion::plan).  For random circuits of every
// gate kind, with controls, it checks the compiled program's structure
// (indices inside their arrays, tile and phase layouts, the LDS swizzle) and
// that running it on a random state equals the oracle's plain apply.  Built
// with address/undefined-behaviour sanitizers by tests/test_host_compiler.py.
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../../oracle/qvnt_oracle.h"
#include "../../qvnt_amd/csrc/fusion.hpp"

using namespace qvnt::fusion;
typedef std::complex<double> C;

static int failures = 0;
#define CHECK(c)                                                                     \
    do {                                                                             \
        if (!(c)) {                                                                  \
            std::fprintf(stderr, "CHECK failed %s:%d %s\n", __FILE__, __LINE__, #c); \
            std::exit(1);                                                            \
        }                                                                            \
    } while (0)

static uint32_t swz(uint32_t e) { return e ^ ((e >> 4) & 15u); }
static C cf(const Program& P, uint32_t off, int k) {
    CHECK(off + 2 * k + 1 < P.coef.size());
    return C(P.coef[off + 2 * k], P.coef[off + 2 * k + 1]);
}

// T2's and PH's coefficient, carried in the op (FOp::ga: real, gb: imaginary)
static C inline_coef(const FOp& op) {
    double re, im;
    std::memcpy(&re, &op.ga, sizeof re);
    std::memcpy(&im, &op.gb, sizeof im);
    return C(re, im);
}

// One register phase of one thread: the ops of fused.hip's run_phase.

static void run_phase(C* v, const FPhase& ph, const Program& P, uint32_t tb, uint64_t tile) {
    CHECK(ph.op_begin <= ph.op_end && ph.op_end <= P.ops.size());
    // the kernel runs a phase's G ops after all others: they must come last
    for (uint32_t o = ph.op_begin; o + 1 < ph.op_end; ++o)
        CHECK(P.ops[o].type != OP_G || P.ops[o + 1].type == OP_G);
    C scal = 1.0;
    for (uint32_t o = ph.op_begin; o < ph.op_end; ++o) {
        const FOp& op = P.ops[o];
        if ((tb & op.lctrl) != op.lctrl || (tile & op.gctrl) != op.gctrl) continue;
        if (op.type == OP_G) {
            CHECK(!op.la && !op.lb && !op.lctrl);
            const uint32_t k = ((tile & op.ga) != 0 ? 1u : 0u) | ((tile & op.gb) != 0 ? 2u : 0u);
            for (int i = 1; i < kRegs; ++i) CHECK(cf(P, op.coef, k * kRegs + i) == cf(P, op.coef, k * kRegs));
            scal = cf(P, op.coef, k * kRegs) * scal;
        } else if (op.type == OP_D) {
            const uint32_t k = ((((tb & op.la) | (tile & op.ga)) != 0) ? 1u : 0u) |
                               ((((tb & op.lb) | (tile & op.gb)) != 0) ? 2u : 0u);
            const bool ea = op.la || op.ga, eb = op.lb || op.gb;
            if (op.dmode == 1) {
                CHECK(ea != eb);
                const uint32_t var = ea ? 1 : 2;
                for (int i = 0; i < kRegs; ++i) v[i] = cf(P, op.coef, i) * v[i];
                if (k & var)
                    for (int i = 0; i < kRegs; ++i) v[i] = cf(P, op.coef, var * kRegs + i) * v[i];
            } else {
                for (int i = 0; i < kRegs; ++i) v[i] = cf(P, op.coef, k * kRegs + i) * v[i];
            }
        } else if (op.type == OP_PH) {
            CHECK(op.sel < (uint32_t)kRegBits && !op.la && !op.lb);
            const C g = inline_coef(op);
            for (int i = 0; i < kRegs; ++i)
                if ((i >> op.sel) & 1) v[i] = g * v[i];
        } else if (op.type == OP_S) {
            // sel: which of the two diagonals the kernel applies
            CHECK(op.sel <= 3);
            if (op.sel & 1u)
                for (int i = 0; i < kRegs; ++i) v[i] = cf(P, op.coef, i) * v[i];
            for (int j = 0; j < kRegBits; ++j) {
                CHECK(cf(P, op.coef, kRegs + j).imag() == 0.0);
                const double a = cf(P, op.coef, kRegs + j).real();
                for (int i = 0; i < kRegs; ++i) {
                    if (i & (1 << j)) continue;
                    const C x0 = v[i], x1 = v[i | (1 << j)];
                    v[i] = x0 - a * x1;
                    v[i | (1 << j)] = x1 + a * x0;
                }
            }
            if (op.sel & 2u)
                for (int i = 0; i < kRegs; ++i) v[i] = cf(P, op.coef, kRegs + kRegBits + i) * v[i];
        } else if (op.type == OP_M2 || op.type == OP_A2 || op.type == OP_X2 || op.type == OP_T2) {
            const int J = (int)op.sel;
            CHECK(J < kRegBits);
            for (int i = 0; i < kRegs; ++i) {
                if (i & (1 << J)) continue;
                if (op.pc != 0xFFFFu && !((op.pc >> i) & 1u)) continue;
                const C x0 = v[i], x1 = v[i | (1 << J)];
                if (op.type == OP_M2) {
                    v[i] = cf(P, op.coef, 0) * x0 + cf(P, op.coef, 1) * x1;
                    v[i | (1 << J)] = cf(P, op.coef, 2) * x0 + cf(P, op.coef, 3) * x1;
                } else if (op.type == OP_A2) {
                    v[i] = cf(P, op.coef, 0) * x1;
                    v[i | (1 << J)] = cf(P, op.coef, 1) * x0;
                } else if (op.type == OP_T2) {
                    const double a = inline_coef(op).real();
                    CHECK(inline_coef(op).imag() == 0.0);
                    v[i] = x0 - a * x1;
                    v[i | (1 << J)] = x1 + a * x0;
                } else {
                    v[i] = x1;
                    v[i | (1 << J)] = x0;
                }
            }
        } else {
            CHECK(op.type == OP_M4 || op.type == OP_P4 || op.type == OP_P4R);
            const int j0 = (int)op.sel / 4, j1 = (int)op.sel % 4;
            CHECK(j0 < j1 && j1 < kRegBits);
            for (int i = 0; i < kRegs; ++i) {
                if (i & ((1 << j0) | (1 << j1))) continue;
                if (!((op.pc >> i) & 1u)) continue;
                const int idx[4] = {i, i | (1 << j0), i | (1 << j1), i | (1 << j0) | (1 << j1)};
                C x[4], y[4];
                for (int k = 0; k < 4; ++k) x[k] = v[idx[k]];
                for (int r = 0; r < 4; ++r) {
                    if (op.type == OP_M4) {
                        y[r] = 0.0;
                        for (int k = 0; k < 4; ++k) y[r] += cf(P, op.coef, 4 * r + k) * x[k];
                    } else if (op.type == OP_P4R) {
                        const C k = cf(P, op.coef, r);
                        y[r] = k.real() * x[r] + C(0.0, k.imag()) * x[r ^ 3];
                    } else {
                        y[r] = cf(P, op.coef, 2 * r) * x[r] + cf(P, op.coef, 2 * r + 1) * x[r ^ 3];
                    }
                }
                for (int r = 0; r < 4; ++r) v[idx[r]] = y[r];
            }
        }
    }
    for (int i = 0; i < kRegs; ++i) v[i] = scal * v[i];
}

// One pass over the state: every tile in turn (fused.hip: fused_tile).
// The tile's last read-out from LDS into the direct layout, through the
// pass's qubit permutation when it has one (the kernel's read_out); a
// permuted read-out must visit every LDS element exactly once.
static void read_out(std::vector<C>& V, const std::vector<C>& lds, const FPass& pass) {
    const uint32_t t = pass.t, nth = pass.nthreads;
    if (!pass.permuted) {
        for (uint32_t tid = 0; tid < nth; ++tid)
            for (int j = 0; j < kRegs; ++j) V[tid * kRegs + j] = lds[swz(tid) ^ swz(j * nth)];
        return;
    }
    CHECK(!pass.direct_last);
    std::vector<char> seen(1u << t, 0);
    for (uint32_t tid = 0; tid < nth; ++tid) {
        uint32_t s = 0;
        for (uint32_t k = 0; k + kRegBits < t; ++k)
            if (tid >> k & 1) s ^= pass.pimg[k];
        for (int j = 0; j < kRegs; ++j) {
            const uint32_t off = s ^ pass.pout[j];
            CHECK(off < (1u << t) && !seen[off]);
            seen[off] = 1;
            V[tid * kRegs + j] = lds[off];
        }
    }
}

static void run_pass(C* psi, uint64_t len, const Program& P, const FPass& pass) {
    const uint32_t t = pass.t, nth = pass.nthreads;
    CHECK(t >= 4 && t <= (uint32_t)kMaxTile && nth == (1u << (t - kRegBits)));
    for (uint32_t k = 1; k < t; ++k) CHECK(pass.tpos[k - 1] < pass.tpos[k]);
    for (uint32_t k = 0; k < 16; ++k) CHECK(((pass.tpack[k / 4] >> (8 * (k % 4))) & 255u) == (k < t ? pass.tpos[k] : 0u));
    CHECK((1ull << pass.tpos[t - 1]) < len);
    CHECK(pass.phase_begin < pass.phase_end && pass.phase_end <= P.phases.size());
    std::vector<C> lds(1u << t), V(nth * kRegs);
    for (uint64_t blk = 0; blk < (len >> t); ++blk) {
        uint64_t tile = blk;
        for (uint32_t k = 0; k < t; ++k) {
            const uint64_t lo = tile & ((1ull << pass.tpos[k]) - 1u);
            tile = ((tile ^ lo) << 1) | lo;
        }
        std::vector<uint64_t> lo_off(nth);
        for (uint32_t tid = 0; tid < nth; ++tid) {
            uint64_t lo = 0;
            for (uint32_t k = 0; k + kRegBits < t; ++k) lo |= (uint64_t)((tid >> k) & 1u) << pass.tpos[k];
            lo_off[tid] = lo;
            for (int j = 0; j < kRegs; ++j) {
                const uint64_t g = tile | lo | pass.hoff[j];
                CHECK(g < len);
                V[tid * kRegs + j] = psi[g];
            }
        }
        for (uint32_t p = pass.phase_begin; p < pass.phase_end; ++p) {
            const FPhase& ph = P.phases[p];
            uint32_t roff[kRegs];
            for (int k = 1; k < kRegBits; ++k) CHECK(ph.rpos[k - 1] < ph.rpos[k]);
            CHECK(ph.rpos[kRegBits - 1] < t);
            for (int i = 0; i < kRegs; ++i) {
                uint32_t off = 0;
                for (int b = 0; b < kRegBits; ++b)
                    if (i >> b & 1) off |= 1u << ph.rpos[b];
                roff[i] = off;
                CHECK(ph.sroff[i] == swz(off));
            }
            const bool din = (p == pass.phase_begin) && pass.direct_first;
            const bool dout = (p + 1 == pass.phase_end) && pass.direct_last;
            if (!din && p == pass.phase_begin)
                for (uint32_t tid = 0; tid < nth; ++tid)
                    for (int j = 0; j < kRegs; ++j) lds[swz(tid) ^ swz(j * nth)] = V[tid * kRegs + j];
            if (ph.dense) {
                CHECK(!din && !dout && t >= (uint32_t)kMinDenseTile && ph.op_begin == ph.op_end);
                CHECK(ph.ucoef + 2 * kRegs * kRegs <= P.coef.size());
                for (uint32_t tid = 0; tid < nth; ++tid) {
                    uint32_t tb = tid;
                    for (int k = 0; k < kRegBits; ++k) {
                        const uint32_t lo = tb & ((1u << ph.rpos[k]) - 1u);
                        tb = ((tb ^ lo) << 1) | lo;
                    }
                    C x[kRegs];
                    for (int j = 0; j < kRegs; ++j) x[j] = lds.at(swz(tb) ^ swz(roff[j]));
                    for (int m = 0; m < kRegs; ++m) {
                        C acc = 0.0;
                        for (int j = 0; j < kRegs; ++j) acc += cf(P, ph.ucoef, kRegs * m + j) * x[j];
                        lds.at(swz(tb) ^ swz(roff[m])) = acc;
                    }
                }
                if (p + 1 == pass.phase_end) read_out(V, lds, pass);
                continue;
            }
            for (uint32_t tid = 0; tid < nth; ++tid) {
                uint32_t tb = tid;
                for (int k = 0; k < kRegBits; ++k) {
                    const uint32_t lo = tb & ((1u << ph.rpos[k]) - 1u);
                    tb = ((tb ^ lo) << 1) | lo;
                }
                C* v = &V[tid * kRegs];
                if (!din)
                    for (int i = 0; i < kRegs; ++i) v[i] = lds.at(swz(tb) ^ ph.sroff[i]);
                else
                    for (int i = 0; i < kRegs; ++i) CHECK((tb | roff[i]) == tid + i * nth);
                run_phase(v, ph, P, tb, tile);
                if (!dout)
                    for (int i = 0; i < kRegs; ++i) lds.at(swz(tb) ^ ph.sroff[i]) = v[i];
                else
                    for (int i = 0; i < kRegs; ++i) CHECK((tb | roff[i]) == tid + i * nth);
            }
            if (!dout && p + 1 == pass.phase_end) read_out(V, lds, pass);
        }
        for (uint32_t tid = 0; tid < nth; ++tid)
            for (int j = 0; j < kRegs; ++j) psi[tile | lo_off[tid] | pass.hoff[j]] = V[tid * kRegs + j];
    }
}

// Random gates of every kind through the oracle's builders, with controls.
static std::vector<qvnt_gate> random_circuit(std::mt19937_64& g, unsigned n, int count) {
    std::vector<qvnt_gate> out;
    std::uniform_real_distribution<double> ang(0.0, 6.283185307179586);
    auto bit = [&](uint64_t avoid) {
        for (;;) {
            const uint64_t b = 1ull << (g() % n);
            if (!(b & avoid)) return b;
        }
    };
    qvnt_gate buf[64];
    for (int k = 0; k < count; ++k) {
        const uint64_t a = bit(0), b = bit(a);
        long m = 0;
        switch (g() % 18) {
        case 0: m = orc_op_x(a | (g() % 3 == 0 ? b : 0), buf); break;
        case 1: m = orc_op_y(a, buf); break;
        case 2: m = orc_op_z(a | b, buf); break;
        case 3: m = orc_op_s(a, buf); break;
        case 4: m = orc_op_t(a, buf); break;
        case 5: m = orc_op_rx(ang(g), a, buf); break;
        case 6: m = orc_op_ry(ang(g), a, buf); break;
        case 7: m = orc_op_rz(ang(g), a, buf); break;
        case 8: m = orc_op_rxx(ang(g), a | b, buf); break;
        case 9: m = orc_op_ryy(ang(g), a | b, buf); break;
        case 10: m = orc_op_rzz(ang(g), a | b, buf); break;
        case 11: m = orc_op_h(a | b, buf); break;
        case 12: m = orc_op_swap(a | b, buf); break;
        case 13: m = orc_op_i_swap(a | b, buf); break;
        case 14: m = orc_op_sqrt_swap(a | b, buf); break;
        case 15: m = orc_op_sqrt_i_swap(a | b, buf); break;
        case 16: m = orc_op_phi(ang(g), a, buf); break;
        default: m = orc_op_u3(ang(g), ang(g), ang(g), a, buf); break;
        }
        CHECK(m > 0 && m < 64);
        if (n >= 3 && g() % 3 == 0) {
            const uint64_t c = bit(orc_op_act_on(buf, (size_t)m));
            CHECK(orc_op_c(buf, (size_t)m, c) == 0);
        }
        if (g() % 11 == 0) orc_op_dgr(buf, (size_t)m);
        out.insert(out.end(), buf, buf + m);
    }
    return out;
}

// Layers of u3 on every qubit, rzz on neighbours (a ring, so the last
// edge wraps) and one controlled x: every phase of such a pass is a
// structured block, and the diagonals cross tiles and phases.
static std::vector<qvnt_gate> layered_circuit(std::mt19937_64& g, unsigned n, int layers) {
    std::vector<qvnt_gate> out;
    std::uniform_real_distribution<double> ang(0.0, 6.283185307179586);
    qvnt_gate buf[64];
    for (int l = 0; l < layers; ++l) {
        for (unsigned q = 0; q < n; ++q) {
            const long m = orc_op_u3(ang(g), ang(g), ang(g), 1ull << q, buf);
            out.insert(out.end(), buf, buf + m);
        }
        for (unsigned q = l & 1; q < n; q += 2) {
            const long m = orc_op_rzz(ang(g), (1ull << q) | (1ull << ((q + 1) % n)), buf);
            out.insert(out.end(), buf, buf + m);
        }
        const unsigned c = (unsigned)(g() % n), t = (c + 1 + (unsigned)(g() % (n - 1))) % n;
        long m = orc_op_x(1ull << t, buf);
        CHECK(orc_op_c(buf, (size_t)m, 1ull << c) == 0);
        out.insert(out.end(), buf, buf + m);
    }
    return out;
}

// One brickwork layer on 19 qubits whose cx runs from qubit 18 to qubit 3.
// In tiles of 11 (rows 0..2 and 8 free qubits a pass) its 16 other qubits
// fill exactly two passes; formed from the queue's start they put qubit
// 18's u3 in the second pass and qubit 3's in the first, and the cx then
// needs qubit 3 in a third pass.  Formed from the queue's end, two passes.
static std::vector<qvnt_gate> cx_high_layer(std::mt19937_64& g, unsigned n) {
    std::vector<qvnt_gate> out;
    std::uniform_real_distribution<double> ang(0.0, 6.283185307179586);
    qvnt_gate buf[64];
    for (unsigned q = 0; q < n; ++q) {
        const long m = orc_op_u3(ang(g), ang(g), ang(g), 1ull << q, buf);
        out.insert(out.end(), buf, buf + m);
    }
    for (unsigned q = 1; q + 1 < n; q += 2) {
        const long m = orc_op_rzz(ang(g), (1ull << q) | (1ull << (q + 1)), buf);
        out.insert(out.end(), buf, buf + m);
    }
    long m = orc_op_x(1ull << 3, buf);
    CHECK(orc_op_c(buf, (size_t)m, 1ull << (n - 1)) == 0);
    out.insert(out.end(), buf, buf + m);
    return out;
}

// Programs equal field by field (coefficients bit for bit).
static bool same_steps(const std::vector<Step>& a, const std::vector<Step>& b) {
    if (a.size() != b.size()) return false;
    for (size_t k = 0; k < a.size(); ++k) {
        const Step &x = a[k], &y = b[k];
        if (x.plain != y.plain) return false;
        if (x.plain) {
            if (std::memcmp(&x.gate, &y.gate, sizeof x.gate) != 0) return false;
            continue;
        }
        const Program &p = x.prog, &q = y.prog;
        if (p.passes.size() != q.passes.size() || p.phases.size() != q.phases.size() ||
            p.ops.size() != q.ops.size() || p.coef.size() != q.coef.size())
            return false;
        for (size_t i = 0; i < p.passes.size(); ++i) {
            const FPass &u = p.passes[i], &v = q.passes[i];
            if (u.t != v.t || u.nthreads != v.nthreads || u.phase_begin != v.phase_begin ||
                u.phase_end != v.phase_end || u.direct_first != v.direct_first || u.direct_last != v.direct_last ||
                std::memcmp(u.tpos, v.tpos, sizeof u.tpos) != 0 || std::memcmp(u.hoff, v.hoff, sizeof u.hoff) != 0 ||
                u.permuted != v.permuted || std::memcmp(u.pimg, v.pimg, sizeof u.pimg) != 0 ||
                std::memcmp(u.pout, v.pout, sizeof u.pout) != 0 || std::memcmp(u.tpack, v.tpack, sizeof u.tpack) != 0)
                return false;
        }
        if (std::memcmp(p.phases.data(), q.phases.data(), p.phases.size() * sizeof(FPhase)) != 0 ||
            std::memcmp(p.ops.data(), q.ops.data(), p.ops.size() * sizeof(FOp)) != 0 ||
            std::memcmp(p.coef.data(), q.coef.data(), p.coef.size() * sizeof(double)) != 0)
            return false;
    }
    return true;
}

// structured ops run, by form (sel: which of the two diagonals they apply)
static unsigned s_forms[4] = {0, 0, 0, 0};

static void check_case(unsigned n, unsigned tq, unsigned dense, uint64_t seed, bool layered = false,
                       bool autotile = false, unsigned amp_bytes = 16, unsigned want_passes = 0) {
    std::mt19937_64 g(seed);
    const std::vector<qvnt_gate> gates = want_passes ? cx_high_layer(g, n)
                                         : layered ? layered_circuit(g, n, 5) : random_circuit(g, n, 160);
    std::vector<Step> steps;
    if (autotile) {
        // the tile size chosen by the run-time model (tq reports it)
        tq = plan_auto(gates.data(), gates.size(), n, true, steps, layered, true);
        if (tq != 10 && tq != 11 && tq != 12) {
            std::printf("n=%u: plan_auto kept tile %u\n", n, tq);
            ++failures;
        }
    } else {
        // layered: single-gate passes go plain; amp_bytes 8: the FP32 tile rows
        plan(gates.data(), gates.size(), n, tq, steps, dense, layered, false, amp_bytes);
    }
    if (!autotile) {
        // passes compiled on several threads and concatenated: the same program
        std::vector<Step> par;
        plan(gates.data(), gates.size(), n, tq, par, dense, layered, true, amp_bytes);
        if (!same_steps(steps, par)) {
            std::printf("n=%u tq=%u dense=%u: the parallel compile differs\n", n, tq, dense);
            ++failures;
        }
    }
    orc_reg* ref = orc_reg_with_state(n, 0);
    const uint64_t len = orc_reg_len(ref);
    std::normal_distribution<double> nd;
    std::vector<C> psi(len);
    for (auto& z : psi) z = C(nd(g), nd(g));
    std::memcpy(orc_reg_psi(ref), psi.data(), len * sizeof(C));
    orc_reg_apply(ref, gates.data(), gates.size());
    unsigned passes = 0, dense_phases = 0, split = 0, plain = 0, permuted = 0;
    for (const Step& st : steps) {
        if (st.plain) {
            ++plain;
            orc_reg* one = orc_reg_with_state(n, 0);  // a gate the fused kernel does not take
            std::memcpy(orc_reg_psi(one), psi.data(), len * sizeof(C));
            orc_reg_apply(one, &st.gate, 1);
            std::memcpy(psi.data(), orc_reg_psi(one), len * sizeof(C));
            orc_reg_free(one);
            continue;
        }
        for (const FPhase& ph : st.prog.phases) dense_phases += ph.dense;
        for (const FOp& op : st.prog.ops) {
            split += op.type == OP_T2 || op.type == OP_S;
            if (op.type == OP_S && op.sel < 4) ++s_forms[op.sel];
        }
        for (const FPass& fp : st.prog.passes) {
            run_pass(psi.data(), len, st.prog, fp);
            ++passes;
            permuted += fp.permuted;
        }
    }
    if (want_passes && passes + plain != want_passes) {
        std::printf("n=%u tq=%u: %u steps, want %u\n", n, tq, passes + plain, want_passes);
        ++failures;
    }
    const C* want = reinterpret_cast<const C*>(orc_reg_psi(ref));
    double err = 0.0;
    for (uint64_t i = 0; i < len; ++i) err = std::max(err, std::abs(psi[i] - want[i]));
    std::printf("n=%u tq=%u dense=%u: %zu gates, %u passes (%u permuted), %u plain steps, %u dense phases, "
                "%u split/structured ops, max |d| %.2e\n", n, tq, dense, gates.size(), passes, permuted, plain,
                dense_phases, split, err);
    if (!(err <= 1e-12)) ++failures;
    orc_reg_free(ref);
}

int main() {
    const unsigned cases[][3] = {{6, 6, 0}, {7, 6, 0}, {9, 8, 0}, {10, 10, 2}, {11, 9, 0}, {12, 12, 3},
                                 {13, 12, 2}, {13, 13, 0}, {14, 11, 0}, {14, 13, 3}, {15, 12, 2}};
    uint64_t seed = 1;
    for (const auto& c : cases) check_case(c[0], c[1], c[2], seed++);
    const unsigned layered[][3] = {{10, 10, 0}, {13, 12, 0}, {14, 11, 0}, {15, 13, 0}, {16, 12, 3}};
    for (const auto& c : layered) check_case(c[0], c[1], c[2], seed++, true);
    for (unsigned n : {12u, 14u, 16u}) check_case(n, 0, 0, seed++, n == 16, true);
    for (unsigned n : {9u, 13u, 15u}) check_case(n, 11, n == 13 ? 2 : 0, seed++, false, false, 8);
    // passes formed from the queue's end where that needs fewer
    check_case(19, 11, 0, seed++, false, false, 16, 2);
    // every form of the structured op is exercised
    std::printf("structured ops by form: %u %u %u %u\n", s_forms[0], s_forms[1], s_forms[2], s_forms[3]);
    for (unsigned c : s_forms)
        if (!c) ++failures;
    std::printf("%d failure(s)\n", failures);
    return failures ? 1 : 0;
}
