// Host-side compiler of the fused form: lowering, one-qubit merging, pass
// and phase scheduling, encoding (declarations and overview in fusion.hpp).
#include "fusion.hpp"

#include <algorithm>
#include <atomic>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <thread>

namespace qvnt {
namespace fusion {
namespace {

using cd = std::complex<double>;

struct Prim {
    uint8_t type;          // OP_M2 / OP_M4 / OP_D
    int q0 = -1, q1 = -1;  // M2: q0; M4: q0 (matrix bit 0), q1 (matrix bit 1); D: table bits
    long src = -1;         // index of the SingleOp it comes from (-1: several)
    uint64_t ctrl = 0;
    cd m[16];              // M2 2x2, M4 4x4 (row-major), D table[4]
    uint64_t targets() const {
        if (type == OP_D) return 0;
        return (1ull << q0) | (q1 >= 0 ? (1ull << q1) : 0);
    }
    uint64_t bits() const {
        uint64_t b = ctrl | (1ull << q0);
        if (q1 >= 0) b |= 1ull << q1;
        return b;
    }
};

// An uncontrolled swap of qubits q0 and q1, taken out of the queue: the
// compiler relabels the qubits of every later primitive instead of moving
// amplitudes (see relabel_swaps).
constexpr uint8_t kRelabel = 0xF0;

// The primitive with every qubit q moved to to[q].
void remap(Prim& p, const int* to) {
    p.q0 = to[p.q0];
    if (p.q1 >= 0) p.q1 = to[p.q1];
    uint64_t c = 0;
    for (uint64_t m = p.ctrl; m; m &= m - 1) c |= 1ull << to[ctz(m)];
    p.ctrl = c;
}

constexpr double kR = 0.70710678118654752440;

Prim m2(int q, cd a, cd b, cd c, cd d, uint64_t ctrl) {
    Prim p;
    p.type = OP_M2;
    p.q0 = q;
    p.ctrl = ctrl;
    p.m[0] = a; p.m[1] = b; p.m[2] = c; p.m[3] = d;
    return p;
}
Prim diag1(int q, cd e0, cd e1, uint64_t ctrl) {
    Prim p;
    p.type = OP_D;
    p.q0 = q;
    p.ctrl = ctrl;
    p.m[0] = e0; p.m[1] = e1; p.m[2] = e0; p.m[3] = e1;
    return p;
}
Prim diag2(int q0, int q1, cd t0, cd t1, cd t2, cd t3, uint64_t ctrl) {
    Prim p;
    p.type = OP_D;
    p.q0 = q0;
    p.q1 = q1;
    p.ctrl = ctrl;
    p.m[0] = t0; p.m[1] = t1; p.m[2] = t2; p.m[3] = t3;
    return p;
}
// A two-qubit phase table f(a, b) (t[a | 2b]) as three factors that
// commute: f = t0 * (t1/t0)^a * (t2/t0)^b * gamma^(ab), gamma = t3 t0 /
// (t1 t2) -- a phase on qubit a, one on qubit b, and a phase on |11>
// (qubit a's phase controlled by b, dropped when it is 1).  The one-qubit
// factors can go to whichever phase holds their qubit in registers; only
// the |11> phase needs both qubits.
void push_diag2(std::vector<Prim>& out, int q0, int q1, cd t0, cd t1, cd t2, cd t3, uint64_t ctrl) {
    out.push_back(diag1(q0, t0, t1, ctrl));
    out.push_back(diag1(q1, 1.0, t2 / t0, ctrl));
    const cd g = t3 * t0 / (t1 * t2);
    if (g != cd(1.0, 0.0)) out.push_back(diag1(q0, 1.0, g, ctrl | (1ull << q1)));
}

Prim m4(int q0, int q1, const cd* u, uint64_t ctrl) {
    Prim p;
    p.type = OP_M4;
    p.q0 = q0;
    p.q1 = q1;
    p.ctrl = ctrl;
    for (int k = 0; k < 16; ++k) p.m[k] = u[k];
    return p;
}

// Matrices of every reference atomic op from its stored parameters: the
// same linear map as the gather formulas of atomic/<gate>.rs (the plain
// kernels), in matrix form.  Returns false if the gate cannot be fused.
bool lower(const qvnt_gate& g, uint64_t lmask, std::vector<Prim>& out) {
    const uint64_t ctrl = g.ctrl;
    const cd c(g.re, 0.0), s(g.im, 0.0), I(0.0, 1.0);
    const cd e(g.re, g.im);
    auto bitlist = [](uint64_t m) {
        std::vector<int> b;
        for (; m; m &= m - 1) b.push_back((int)ctz(m));
        return b;
    };
    switch (g.kind) {
    case QVNT_X:
        for (int q : bitlist(g.a)) out.push_back(m2(q, 0.0, 1.0, 1.0, 0.0, ctrl));
        return true;
    case QVNT_Y:  // Y^{(x)k}: one Y per bit
        for (int q : bitlist(g.a)) out.push_back(m2(q, 0.0, -I, I, 0.0, ctrl));
        return true;
    case QVNT_Z:
        for (int q : bitlist(g.a & lmask)) out.push_back(diag1(q, 1.0, -1.0, ctrl));
        return true;
    case QVNT_S:
        for (int q : bitlist(g.a & lmask)) out.push_back(diag1(q, 1.0, g.dagger ? -I : I, ctrl));
        return true;
    case QVNT_T: {
        const cd t(kR, g.dagger ? -kR : kR);
        for (int q : bitlist(g.a & lmask)) out.push_back(diag1(q, 1.0, t, ctrl));
        return true;
    }
    case QVNT_RX:
        out.push_back(m2((int)ctz(g.a), c, -I * s, -I * s, c, ctrl));
        return true;
    case QVNT_RY:
        out.push_back(m2((int)ctz(g.a), c, -s, s, c, ctrl));
        return true;
    case QVNT_RZ:
        if ((g.a & lmask) == 0) {  // qubit beyond the buffer: reads as 0, a global phase
            out.push_back(diag1(0, std::conj(e), std::conj(e), ctrl));
            return true;
        }
        out.push_back(diag1((int)ctz(g.a), std::conj(e), e, ctrl));
        return true;
    case QVNT_PHI:
        if ((g.a & lmask) == 0) return true;
        out.push_back(diag1((int)ctz(g.a), 1.0, e, ctrl));
        return true;
    case QVNT_H1:
        out.push_back(m2((int)ctz(g.a), kR, kR, kR, -kR, ctrl));
        return true;
    case QVNT_U1:
        out.push_back(m2((int)ctz(g.a), cd(g.m[0], g.m[1]), cd(g.m[2], g.m[3]), cd(g.m[4], g.m[5]), cd(g.m[6], g.m[7]), ctrl));
        return true;
    case QVNT_RZZ: {
        const std::vector<int> b = bitlist(g.a);
        const std::vector<int> in = bitlist(g.a & lmask);
        if (in.size() == 2) {
            push_diag2(out, b[0], b[1], std::conj(e), e, e, std::conj(e), ctrl);
        } else if (in.size() == 1) {
            out.push_back(diag1(in[0], std::conj(e), e, ctrl));
        } else {
            out.push_back(diag1(0, std::conj(e), std::conj(e), ctrl));
        }
        return true;
    }
    default: break;
    }
    // two-qubit matrices, index bit 0 = lower qubit
    const std::vector<int> b = bitlist(g.kind == QVNT_H2 || g.kind == QVNT_U2 ? (g.a | g.b) : g.a);
    if ((g.kind == QVNT_SWAP || g.kind == QVNT_ISWAP) && ctrl == 0) {
        // iswap = swap * diag(1, w, w, 1) (w = i, or -i daggered): the
        // diagonal is symmetric in the two qubits, so it stays a phase
        // table on them, and the swap becomes a relabeling
        if (g.kind == QVNT_ISWAP) {
            const cd w = g.dagger ? -I : I;
            push_diag2(out, b[0], b[1], 1.0, w, w, 1.0, 0);
        }
        Prim r;
        r.type = kRelabel;
        r.q0 = b[0];
        r.q1 = b[1];
        out.push_back(r);
        return true;
    }
    cd u[16];
    for (int k = 0; k < 16; ++k) u[k] = 0.0;
    auto set = [&](int r, int col, cd v) { u[4 * r + col] = v; };
    switch (g.kind) {
    case QVNT_RXX:
        set(0, 0, c); set(1, 1, c); set(2, 2, c); set(3, 3, c);
        set(0, 3, -I * s); set(1, 2, -I * s); set(2, 1, -I * s); set(3, 0, -I * s);
        break;
    case QVNT_RYY:
        set(0, 0, c); set(1, 1, c); set(2, 2, c); set(3, 3, c);
        set(0, 3, I * s); set(1, 2, -I * s); set(2, 1, -I * s); set(3, 0, I * s);
        break;
    case QVNT_SWAP:
        set(0, 0, 1.0); set(1, 2, 1.0); set(2, 1, 1.0); set(3, 3, 1.0);
        break;
    case QVNT_ISWAP: {
        const cd w = g.dagger ? -I : I;
        set(0, 0, 1.0); set(1, 2, w); set(2, 1, w); set(3, 3, 1.0);
        break;
    }
    case QVNT_SQRT_SWAP: {
        const cd q(0.5, g.dagger ? -0.5 : 0.5);
        set(0, 0, 1.0); set(1, 1, q); set(1, 2, std::conj(q)); set(2, 1, std::conj(q)); set(2, 2, q); set(3, 3, 1.0);
        break;
    }
    case QVNT_SQRT_ISWAP: {
        const cd w(0.0, g.dagger ? -kR : kR);
        set(0, 0, 1.0); set(1, 1, kR); set(1, 2, w); set(2, 1, w); set(2, 2, kR); set(3, 3, 1.0);
        break;
    }
    case QVNT_H2:
        for (int r = 0; r < 4; ++r)
            for (int col = 0; col < 4; ++col) set(r, col, (__builtin_popcount(r & col) & 1) ? -0.5 : 0.5);
        break;
    case QVNT_U2: {
        // reference layout: index bit 0 = a, bit 1 = b (atomic/u2.rs:22-49)
        if (g.ctrl & g.b) return false;
        for (int k = 0; k < 16; ++k) u[k] = cd(g.m[2 * k], g.m[2 * k + 1]);
        out.push_back(m4((int)ctz(g.a), (int)ctz(g.b), u, ctrl));
        return true;
    }
    default:
        return false;
    }
    out.push_back(m4(b[0], b[1], u, ctrl));
    return true;
}

void mat2_mul(const cd* a, const cd* b, cd* out) {  // out = a * b
    cd r[4] = {a[0] * b[0] + a[1] * b[2], a[0] * b[1] + a[1] * b[3], a[2] * b[0] + a[3] * b[2], a[2] * b[1] + a[3] * b[3]};
    for (int k = 0; k < 4; ++k) out[k] = r[k];
}

// Multiplies runs of uncontrolled one-qubit primitives per qubit.
std::vector<Prim> merge_single(const std::vector<Prim>& in) {
    std::vector<Prim> out;
    struct Pend { bool on = false; cd m[4]; long src = -1; };
    std::vector<Pend> pend(64);
    auto flush = [&](int q) {
        if (!pend[q].on) return;
        const cd* m = pend[q].m;
        if (m[1] == 0.0 && m[2] == 0.0) out.push_back(diag1(q, m[0], m[3], 0));
        else out.push_back(m2(q, m[0], m[1], m[2], m[3], 0));
        out.back().src = pend[q].src;
        pend[q].on = false;
    };
    for (const Prim& p : in) {
        const bool single = p.ctrl == 0 && p.q1 < 0 && (p.type == OP_M2 || p.type == OP_D);
        if (single) {
            cd u[4];
            if (p.type == OP_M2) { for (int k = 0; k < 4; ++k) u[k] = p.m[k]; }
            else { u[0] = p.m[0]; u[1] = 0.0; u[2] = 0.0; u[3] = p.m[1]; }
            Pend& pq = pend[p.q0];
            if (pq.on) {
                mat2_mul(u, pq.m, pq.m);
                if (pq.src != p.src) pq.src = -1;
            } else {
                for (int k = 0; k < 4; ++k) pq.m[k] = u[k];
                pq.on = true;
                pq.src = p.src;
            }
            continue;
        }
        for (uint64_t b = p.bits(); b; b &= b - 1) flush((int)ctz(b));
        out.push_back(p);
    }
    for (int q = 0; q < 64; ++q) flush(q);
    return out;
}

struct PassPlan {
    uint64_t T = 0;
    std::vector<Prim> prims;
    // the pass ends by moving the amplitude of index bit x (x in T) to bit
    // sigma[x]: a permutation of T's qubits, free in the kernel (the tile's
    // last LDS read-out is permuted)
    bool permuted = false;
    int8_t sigma[64];
};

// The primitives left behind so far in a scan of the queue.  A later
// primitive may overtake them only if it commutes with each: diagonals
// commute with one another, a diagonal and a matrix primitive commute
// when the matrix targets none of the diagonal's qubits, and two matrix
// primitives are taken to commute only on disjoint qubits.
struct Left {
    uint64_t diag_bits = 0, mat_bits = 0, mat_targets = 0;
    bool blocks(const Prim& p) const {
        if (p.type == OP_D) return (p.bits() & mat_targets) != 0;
        return (p.targets() & diag_bits) != 0 || (p.bits() & mat_bits) != 0;
    }
    void add(const Prim& p) {
        if (p.type == OP_D) diag_bits |= p.bits();
        else { mat_bits |= p.bits(); mat_targets |= p.targets(); }
    }
};

// Qubits 0..row_bits-1 belong to every pass, so each tile row is that many
// consecutive amplitudes in HBM: one 128-byte cache line (3 bits in FP64, 4
// in FP32).  A shorter row would leave the rest of each line to another
// tile, fetched again later.  QVNT_ROW_BITS (1..6) overrides it for tuning.
unsigned row_bits(unsigned amp_bytes) {
    static const int env = [] {
        const char* s = std::getenv("QVNT_ROW_BITS");
        const int v = s ? std::atoi(s) : 0;
        return v >= 1 && v <= 6 ? v : 0;
    }();
    if (env) return (unsigned)env;
    return amp_bytes <= 8 ? 4u : 3u;
}

// Greedy pass formation (see the file comment).  A second scan with the
// pass's final qubit set T leaves for a later pass every diagonal that
// touches no qubit of T (in this pass it would only scale whole tiles; in
// a later one its qubits may be registers), provided the pass keeps some
// matrix primitive, so every pass makes progress.
//
// want (relabeled queues, see relabel_swaps): the primitives address the
// qubit layout in which qubit f's amplitudes belong at index bit want[f]
// once the queue is done.  Each pass then also permutes the qubits of its
// set T (PassPlan::sigma), moving every amplitude bit whose home is in T
// there; T's free slots are filled with qubits that let it move more; the
// primitives left are relabeled to follow; and passes without primitives
// finish what is still out of place.
std::vector<PassPlan> schedule_passes(std::vector<Prim> work, unsigned ns, unsigned tq, unsigned rows,
                                      const std::vector<int>* want) {
    std::vector<PassPlan> passes;
    const unsigned rb = ns < rows ? ns : rows;
    const uint64_t forced = (1ull << rb) - 1u;
    // the queue holds indices into `work`, in circuit order
    std::vector<uint32_t> rest(work.size()), leave, take;
    for (size_t k = 0; k < work.size(); ++k) rest[k] = (uint32_t)k;
    // does a matrix primitive after queue position i target one of `bits`?
    auto later_meets = [&](size_t i, uint64_t bits) {
        for (size_t k = i + 1; k < rest.size(); ++k) {
            const Prim& q = work[rest[k]];
            if (q.type != OP_D && (q.targets() & bits)) return true;
        }
        return false;
    };
    // layout: at[x] = the qubit of `work`'s original layout whose amplitude
    // bit now sits at index bit x; home(x) = where that bit belongs
    std::vector<int> at(ns), where(ns);
    for (unsigned q = 0; q < ns; ++q) at[q] = where[q] = (int)q;
    auto home = [&](int x) { return want ? (*want)[(size_t)at[(size_t)x]] : x; };
    auto out_of_place = [&]() {
        for (unsigned x = 0; x < ns; ++x)
            if (home((int)x) != (int)x) return true;
        return false;
    };
    // the bits that belong in T or sit in T, while T has room
    auto pad_moves = [&](uint64_t& T) {
        if (!want) return;
        for (unsigned x = 0; x < ns && popc(T) < tq; ++x)
            if ((T >> x & 1) && home((int)x) != (int)x) T |= 1ull << home((int)x);
        for (unsigned x = 0; x < ns && popc(T) < tq; ++x)
            if (!(T >> x & 1) && (T >> home((int)x) & 1)) T |= 1ull << x;
        for (unsigned x = 0; x < ns && popc(T) + 2 <= tq; ++x)
            if (!(T >> x & 1) && home((int)x) != (int)x && !(T >> home((int)x) & 1))
                T |= (1ull << x) | (1ull << home((int)x));
    };
    // the pass's permutation of T: every bit whose home is in T goes home,
    // the others fill T's remaining bits in order; then the layout and the
    // primitives still queued follow it
    auto permute = [&](PassPlan& pp) {
        for (int x = 0; x < 64; ++x) pp.sigma[x] = (int8_t)x;
        if (!want) return;
        uint64_t placed = 0, taken = 0;
        for (uint64_t m = pp.T; m; m &= m - 1) {
            const int x = (int)ctz(m), h = home(x);
            if (pp.T >> h & 1) {
                pp.sigma[x] = (int8_t)h;
                placed |= 1ull << x;
                taken |= 1ull << h;
            }
        }
        uint64_t src = pp.T & ~placed, dst = pp.T & ~taken;
        for (; src; src &= src - 1, dst &= dst - 1) pp.sigma[ctz(src)] = (int8_t)ctz(dst);
        int to[64];
        for (int x = 0; x < 64; ++x) {
            to[x] = pp.sigma[x];
            if (to[x] != x) pp.permuted = true;
        }
        if (!pp.permuted) return;
        std::vector<int> moved(ns);
        for (unsigned x = 0; x < ns; ++x) moved[(size_t)to[x]] = at[x];
        at.swap(moved);
        for (unsigned x = 0; x < ns; ++x) where[(size_t)at[x]] = (int)x;
        for (uint32_t k : rest) remap(work[k], to);
    };
    while (!rest.empty()) {
        // Once the left-behind matrix primitives target every qubit the
        // queue touches, every later primitive is blocked: a scan stops there.
        uint64_t support = 0;
        for (uint32_t k : rest) support |= work[k].bits();
        PassPlan pp;
        pp.T = forced;
        Left left;
        bool matrix = false;
        leave.clear();
        take.clear();
        size_t i = 0;
        for (; i < rest.size() && (left.mat_targets & support) != support; ++i) {
            const Prim& p = work[rest[i]];
            if (left.blocks(p)) { left.add(p); leave.push_back(rest[i]); continue; }
            const uint64_t nt = pp.T | p.targets();
            if (popc(nt) <= tq) { pp.T = nt; take.push_back(rest[i]); matrix |= p.type != OP_D; }
            else { left.add(p); leave.push_back(rest[i]); }
        }
        leave.insert(leave.end(), rest.begin() + (long)i, rest.end());
        // pad T up to tq: first with bits a permutation can bring home,
        // then with the lowest free qubits (larger, fewer tiles)
        pad_moves(pp.T);
        for (unsigned q = 0; q < ns && popc(pp.T) < tq; ++q) pp.T |= 1ull << q;
        if (matrix) {
            // a deferred diagonal must meet a later matrix target, so some
            // later pass holds its qubits anyway
            take.clear();
            leave.clear();
            left = Left();
            for (i = 0; i < rest.size() && (left.mat_targets & support) != support; ++i) {
                const Prim& p = work[rest[i]];
                if (left.blocks(p) || (p.targets() & ~pp.T) ||
                    (p.type == OP_D && !(p.bits() & pp.T) && later_meets(i, p.bits()))) {
                    left.add(p);
                    leave.push_back(rest[i]);
                    continue;
                }
                take.push_back(rest[i]);
            }
            leave.insert(leave.end(), rest.begin() + (long)i, rest.end());
        }
        pp.prims.reserve(take.size());
        for (uint32_t k : take) pp.prims.push_back(work[k]);
        rest.swap(leave);
        permute(pp);
        passes.push_back(std::move(pp));
    }
    // what is still out of place: passes that only permute, each taking
    // whole cycles of the layout (or as much of one as fits).  Each brings
    // at least one bit home (the first cycle's first two bits fit: the
    // caller relabels only where tq >= rows + 2 or T holds every qubit), so
    // at most ns such passes follow.
    while (want && out_of_place()) {
        PassPlan pp;
        pp.T = forced;
        for (unsigned x0 = 0; x0 < ns && popc(pp.T) < tq; ++x0) {
            if (home((int)x0) == (int)x0 || (pp.T >> x0 & 1)) continue;
            int x = (int)x0;
            do {
                pp.T |= 1ull << x;
                x = home(x);
            } while (x != (int)x0 && popc(pp.T) < tq);
        }
        for (unsigned q = 0; q < ns && popc(pp.T) < tq; ++q) pp.T |= 1ull << q;
        permute(pp);
        passes.push_back(std::move(pp));
    }
    return passes;
}

// The same greedy run from the queue's end: the last pass takes every
// primitive that commutes with all those left after it, and so on
// backwards; the passes are then put in order.  The commutation rules are
// symmetric, so the plan is as valid.  It differs where the queue's order
// of qubits works against it: a controlled gate whose control qubit's
// one-qubit gates fall in a later forward pass than its target's needs the
// target in yet another pass (a brickwork layer's cx with the control above
// the target: 4 passes forward, 3 backward).  Not for relabeled queues
// (their passes also permute the layout, in queue order).
std::vector<PassPlan> schedule_backward(const std::vector<Prim>& work, unsigned ns, unsigned tq, unsigned rows) {
    std::vector<PassPlan> passes = schedule_passes(std::vector<Prim>(work.rbegin(), work.rend()), ns, tq, rows, nullptr);
    std::reverse(passes.begin(), passes.end());
    for (PassPlan& pp : passes) std::reverse(pp.prims.begin(), pp.prims.end());
    return passes;
}

int local_of(const uint32_t* tpos, unsigned t, int q) {
    for (unsigned k = 0; k < t; ++k)
        if ((int)tpos[k] == q) return (int)k;
    return -1;
}

uint32_t add_coef(Program& prog, const cd* m, int n) {
    const uint32_t off = (uint32_t)prog.coef.size();
    for (int k = 0; k < n; ++k) {
        prog.coef.push_back(m[k].real());
        prog.coef.push_back(m[k].imag());
    }
    return off;
}

FPhase make_phase(uint32_t Rl) {
    FPhase ph{};
    int k = 0;
    for (uint32_t m = Rl; m; m &= m - 1) ph.rpos[k++] = ctz(m);
    for (int i = 0; i < kRegs; ++i) {
        uint32_t off = 0;
        for (int b = 0; b < kRegBits; ++b)
            if (i >> b & 1) off |= 1u << ph.rpos[b];
        ph.sroff[i] = lds_swizzle(off);
    }
    return ph;
}

// The 16x16 action on the register index of the ops [op_begin, op_end) of
// a phase whose ops touch register bits only, in the kernel's semantics
// (fused.hip: apply_m2 / apply_m4 / apply_d with table variant 0).
void dense_of(const Program& prog, uint32_t op_begin, uint32_t op_end, cd* U) {
    // real and imaginary parts apart; each product and sum is the one
    // std::complex forms (re = ac - bd, im = ad + bc), without its NaN checks
    double ur[kRegs][kRegs], ui[kRegs][kRegs];
    for (int r = 0; r < kRegs; ++r)
        for (int c = 0; c < kRegs; ++c) {
            ur[r][c] = r == c ? 1.0 : 0.0;
            ui[r][c] = 0.0;
        }
    // U <- G U for an op G that mixes the rows idx[0..m) (ascending) by the
    // m x m block v and leaves every other row alone
    auto mix = [&](const int* idx, int m, const cd* v) {
        double outr[4][kRegs], outi[4][kRegs];
        for (int r = 0; r < m; ++r) {
            for (int c = 0; c < kRegs; ++c) outr[r][c] = outi[r][c] = 0.0;
            for (int k = 0; k < m; ++k) {
                const cd g = v[r * m + k];
                if (g == 0.0) continue;
                const double gr = g.real(), gi = g.imag();
                const double* xr = ur[idx[k]];
                const double* xi = ui[idx[k]];
                for (int c = 0; c < kRegs; ++c) {
                    outr[r][c] += gr * xr[c] - gi * xi[c];
                    outi[r][c] += gr * xi[c] + gi * xr[c];
                }
            }
        }
        for (int r = 0; r < m; ++r)
            for (int c = 0; c < kRegs; ++c) {
                ur[idx[r]][c] = outr[r][c];
                ui[idx[r]][c] = outi[r][c];
            }
    };
    for (uint32_t o = op_begin; o < op_end; ++o) {
        const FOp& op = prog.ops[o];
        auto cf = [&](int k) { return cd(prog.coef[op.coef + 2 * k], prog.coef[op.coef + 2 * k + 1]); };
        if (op.type == OP_D) {
            for (int i = 0; i < kRegs; ++i) {
                const int idx[1] = {i};
                const cd v[1] = {cf(i)};
                mix(idx, 1, v);
            }
        } else if (op.type == OP_M2 || op.type == OP_A2 || op.type == OP_X2) {
            const int J = (int)op.sel;
            const cd one(1.0, 0.0), zero(0.0, 0.0);
            const cd u00 = op.type == OP_M2 ? cf(0) : zero, u11 = op.type == OP_M2 ? cf(3) : zero;
            const cd u01 = op.type == OP_M2 ? cf(1) : op.type == OP_A2 ? cf(0) : one;
            const cd u10 = op.type == OP_M2 ? cf(2) : op.type == OP_A2 ? cf(1) : one;
            const cd v[4] = {u00, u01, u10, u11};
            for (int i = 0; i < kRegs; ++i) {
                if (i & (1 << J)) continue;
                if (op.pc != 0xFFFFu && !((op.pc >> i) & 1u)) continue;
                const int idx[2] = {i, i | (1 << J)};
                mix(idx, 2, v);
            }
        } else {
            const int j0 = (int)op.sel / 4, j1 = (int)op.sel % 4;  // j0 < j1
            cd v[16];
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c) {
                    cd x = 0.0;
                    if (op.type == OP_M4) x = cf(4 * r + c);
                    else if (op.type == OP_P4R && c == r) x = cf(r).real();
                    else if (op.type == OP_P4R && c == (r ^ 3)) x = cd(0.0, cf(r).imag());
                    else if (c == r) x = cf(2 * r);
                    else if (c == (r ^ 3)) x = cf(2 * r + 1);
                    v[4 * r + c] = x;
                }
            for (int i = 0; i < kRegs; ++i) {
                if (i & ((1 << j0) | (1 << j1))) continue;
                if (!((op.pc >> i) & 1u)) continue;
                const int idx[4] = {i, i | (1 << j0), i | (1 << j1), i | (1 << j0) | (1 << j1)};
                mix(idx, 4, v);
            }
        }
    }
    for (int r = 0; r < kRegs; ++r)
        for (int c = 0; c < kRegs; ++c) U[r * kRegs + c] = cd(ur[r][c], ui[r][c]);
}

uint32_t reg_targets(const FOp& op) {
    if (op.type == OP_D) return 0;
    if (op.type == OP_M4 || op.type == OP_P4 || op.type == OP_P4R) return (1u << (op.sel / 4)) | (1u << (op.sel % 4));
    return 1u << op.sel;
}

// Register bits a diagonal's tables depend on.
uint32_t diag_deps(const Program& prog, const FOp& op) {
    uint32_t dep = 0;
    const double* t = prog.coef.data() + op.coef;
    for (int j = 0; j < kRegBits; ++j) {
        const int bit = 1 << j;
        bool differs = false;  // some entry pair i, i | bit of some variant
        for (int k = 0; k < 4 && !differs; ++k)
            for (int i = 0; i < kRegs && !differs; ++i) {
                if (i & bit) continue;
                const double* a = t + 2 * (k * kRegs + i);
                const double* b = a + 2 * bit;
                differs = a[0] != b[0] || a[1] != b[1];
            }
        if (differs) dep |= 1u << j;
    }
    return dep;
}

bool has_ext(const FOp& op) { return op.la || op.lb || op.ga || op.gb; }

// A diagonal on register bits only, without controls outside the
// registers: its variant-0 table is the whole op.
bool plain_diag(const FOp& op) { return op.type == OP_D && !has_ext(op) && !op.lctrl && !op.gctrl; }

// Merges diagonal `d` (later) into diagonal `e` (earlier) if their controls
// and non-register operands allow one table set; true on success.
bool merge_diag(Program& prog, FOp& e, const FOp& d) {
    if (e.lctrl != d.lctrl || e.gctrl != d.gctrl) return false;
    const bool xe = has_ext(e), xd = has_ext(d);
    if (xe && xd && (e.la != d.la || e.lb != d.lb || e.ga != d.ga || e.gb != d.gb)) return false;
    double* te = prog.coef.data() + e.coef;
    const double* td = prog.coef.data() + d.coef;
    // all products first: without operands of its own, e's variant 0 is
    // read for every k
    cd out[4 * kRegs];
    for (int k = 0; k < 4; ++k)
        for (int i = 0; i < kRegs; ++i) {
            const int ke = xe ? k : 0, kd = xd ? k : 0;
            const cd a(te[2 * (ke * kRegs + i)], te[2 * (ke * kRegs + i) + 1]);
            const cd b(td[2 * (kd * kRegs + i)], td[2 * (kd * kRegs + i) + 1]);
            out[k * kRegs + i] = b * a;
        }
    for (int k = 0; k < 4 * kRegs; ++k) {
        te[2 * k] = out[k].real();
        te[2 * k + 1] = out[k].imag();
    }
    if (!xe && xd) {
        e.la = d.la;
        e.lb = d.lb;
        e.ga = d.ga;
        e.gb = d.gb;
    }
    return true;
}

// A diagonal with exactly one non-register operand: its variant slot
// (1 for a, 2 for b) becomes T_var / T_0 (the kernel multiplies by T_0,
// then by the ratio where the operand bit is 1).
void ratio_table(Program& prog, FOp& op, const FPhase& ph) {
    if (op.type != OP_D) return;
    const bool ea = op.la || op.ga, eb = op.lb || op.gb;
    if (ea == eb) return;
    // the operand's bit must be uniform over a wave: outside the tile, or a
    // local position whose thread-index bit (register positions removed) is >= 6
    const uint32_t lm = ea ? op.la : op.lb;
    if (lm) {
        const uint32_t l = ctz(lm);
        uint32_t below = 0;
        for (int k = 0; k < kRegBits; ++k) below += ph.rpos[k] < l ? 1u : 0u;
        if (l - below < 6) return;
    }
    op.dmode = 1;
    const int var = ea ? 1 : 2;
    double* t = prog.coef.data() + op.coef;
    for (int i = 0; i < kRegs; ++i) {
        const cd t0(t[2 * i], t[2 * i + 1]);
        const cd tv(t[2 * (var * kRegs + i)], t[2 * (var * kRegs + i) + 1]);
        const cd q = tv / t0;
        t[2 * (var * kRegs + i)] = q.real();
        t[2 * (var * kRegs + i) + 1] = q.imag();
    }
}

// Diagonals commute with one another and with every matrix op whose target
// register bits their tables do not depend on, so a diagonal may sit
// anywhere between the last earlier and the first later matrix op that
// targets one of its bits (its window of slots; slot s lies before matrix
// op s).  In order of window end, each diagonal joins the first group
// placed inside its window whose table it merges with, or opens a group at
// its window's end; every group is one table.  `deps` caches diag_deps by
// coefficient offset for tables that hold the same values whenever this is
// called with that cache.
using DepCache = std::vector<std::pair<uint32_t, uint32_t>>;
void merge_diagonals(Program& prog, uint32_t begin, bool separate_ext, DepCache& deps) {
    auto dep_of = [&](const FOp& op) {
        for (const auto& e : deps)
            if (e.first == op.coef) return e.second;
        const uint32_t d = diag_deps(prog, op);
        deps.push_back({op.coef, d});
        return d;
    };
    std::vector<FOp> mats;
    struct Diag { FOp op; uint32_t dep, lo, hi, at; };
    std::vector<Diag> diags;
    mats.reserve(prog.ops.size() - begin);
    diags.reserve(prog.ops.size() - begin);
    for (size_t o = begin; o < prog.ops.size(); ++o) {
        const FOp& op = prog.ops[o];
        if (op.type == OP_D) diags.push_back({op, dep_of(op), 0, 0, (uint32_t)mats.size()});
        else mats.push_back(op);
    }
    const uint32_t m = (uint32_t)mats.size();
    // windows: lo = 1 + the last matrix op before the diagonal that targets
    // one of its bits (0 if none), hi = the first one from its place on (m
    // if none); one sweep each way, per register bit (diags are in op order,
    // so their `at` never decreases)
    {
        uint32_t last[kRegBits] = {0, 0, 0, 0};  // 1 + the last matrix op on bit j so far
        size_t di = 0;
        for (uint32_t k = 0; k <= m; ++k) {
            for (; di < diags.size() && diags[di].at == k; ++di) {
                Diag& d = diags[di];
                d.lo = 0;
                for (int j = 0; j < kRegBits; ++j)
                    if ((d.dep >> j) & 1u) d.lo = std::max(d.lo, last[j]);
            }
            if (k < m)
                for (uint32_t t = reg_targets(mats[k]); t; t &= t - 1) last[ctz(t)] = k + 1;
        }
        uint32_t next[kRegBits] = {m, m, m, m};  // the first matrix op on bit j at or after k
        di = diags.size();
        for (uint32_t k = m + 1; k-- > 0;) {
            for (; di > 0 && diags[di - 1].at == k; --di) {
                Diag& d = diags[di - 1];
                d.hi = m;
                for (int j = 0; j < kRegBits; ++j)
                    if ((d.dep >> j) & 1u) d.hi = std::min(d.hi, next[j]);
            }
            if (k > 0)
                for (uint32_t t = reg_targets(mats[k - 1]); t; t &= t - 1) next[ctz(t)] = k - 1;
        }
    }
    std::vector<size_t> order(diags.size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return diags[a].hi < diags[b].hi; });
    struct Group { uint32_t slot; FOp op; };
    std::vector<Group> groups;
    // a T2 that a structured op can take in
    auto foldable = [&](uint32_t s) {
        const FOp& op = mats[s];
        return op.type == OP_T2 && op.pc == 0xFFFFu && !op.lctrl && !op.gctrl;
    };
    // Kept apart, a table with non-register operands or controls must not
    // sit inside a run of foldable T2s (it would split the structured op).
    // It goes to the phase's start where it can, else to a slot a plain
    // group holds (the boundary of a run), else to the phase's end, else
    // to a slot outside any run, else to its window's end.
    auto ext_slot = [&](const Diag& d) {
        if (d.lo == 0) return 0u;
        for (uint32_t s = d.hi + 1; s-- > d.lo;)
            for (const Group& g : groups)
                if (g.slot == s && plain_diag(g.op)) return s;
        if (d.hi == m) return m;
        for (uint32_t s = d.hi + 1; s-- > d.lo;)
            if (!foldable(s - 1) || !foldable(s)) return s;
        return d.hi;
    };
    // with separate_ext, the plain tables are grouped first, so the others
    // can be placed beside them
    for (int round = 0; round < (separate_ext ? 2 : 1); ++round)
        for (size_t k : order) {
            const Diag& d = diags[k];
            if (separate_ext && plain_diag(d.op) != (round == 0)) continue;
            bool merged = false;
            for (Group& g : groups)
                if (g.slot >= d.lo && g.slot <= d.hi && (!separate_ext || plain_diag(g.op) == plain_diag(d.op)) &&
                    merge_diag(prog, g.op, d.op)) {
                    merged = true;
                    break;
                }
            if (!merged) groups.push_back({separate_ext && !plain_diag(d.op) ? ext_slot(d) : d.hi, d.op});
        }
    // within a slot, tables on register bits only sit next to the matrix
    // ops (before the next one, or after the last), where a structured op
    // can take them in
    prog.ops.resize(begin);
    for (uint32_t s = 0; s <= m; ++s) {
        for (int pass = 0; pass < 2; ++pass) {
            const bool plain_now = (s == m) == (pass == 0);
            for (const Group& g : groups)
                if (g.slot == s && plain_diag(g.op) == plain_now) prog.ops.push_back(g.op);
        }
        if (s < m) prog.ops.push_back(mats[s]);
    }
}

// A diagonal on register bit J of the phase, d0 / d1 where the bit is
// 0 / 1, for the register indices in pc (phase 1 elsewhere); all four
// variants alike.
FOp diag_on(Program& prog, uint32_t J, uint32_t pc, cd d0, cd d1) {
    FOp d{};
    d.type = OP_D;
    d.pc = 0xFFFFu;
    cd tab[4 * kRegs];
    for (int k = 0; k < 4; ++k)
        for (int i = 0; i < kRegs; ++i)
            tab[k * kRegs + i] = ((pc >> i) & 1u) ? ((i >> J) & 1 ? d1 : d0) : cd(1.0, 0.0);
    d.coef = add_coef(prog, tab, 4 * kRegs);
    return d;
}

// Splits the 2x2 ops of a phase whose controls are register bits only
// (so the diagonals can fold them in):
//   A2 [[0, a], [b, 0]] = X2 * diag(b, a);
//   M2 = diag(l0, l1) R diag(1, r1) with R = [[c, -s], [s, c]] real,
//   c = |m00|, s = |m10|, l0 = m00 / c, l1 = m10 / s, r1 = m11 / (l1 c),
//   and R = c [[1, -t], [t, 1]] with t = s / c (T2), the scale c going
//   into the left diagonal.  With FMA the T2 outputs x0 - t x1 and
//   x1 + t x0 round once, so a large t (c small) loses nothing once scaled
//   back by c.  An M2 that does not reproduce to 1e-14 (a non-unitary
//   matrix, an overflowing t) or has a zero entry stays whole.
void split_pairs(Program& prog, uint32_t begin) {
    std::vector<FOp> in(prog.ops.begin() + begin, prog.ops.end());
    prog.ops.resize(begin);
    for (const FOp& op : in) {
        if ((op.type != OP_M2 && op.type != OP_A2) || op.lctrl || op.gctrl) {
            prog.ops.push_back(op);
            continue;
        }
        auto cf = [&](int k) { return cd(prog.coef[op.coef + 2 * k], prog.coef[op.coef + 2 * k + 1]); };
        if (op.type == OP_A2) {
            const cd a = cf(0), b = cf(1);
            prog.ops.push_back(diag_on(prog, op.sel, op.pc, b, a));
            FOp x = op;
            x.type = OP_X2;
            x.coef = 0;
            prog.ops.push_back(x);
            continue;
        }
        const cd m00 = cf(0), m01 = cf(1), m10 = cf(2), m11 = cf(3);
        const double c = std::abs(m00), s = std::abs(m10);
        if (!(c > 0.0 && s > 0.0)) {
            prog.ops.push_back(op);
            continue;
        }
        const cd l0 = m00 / c, l1 = m10 / s, r1 = m11 / (l1 * c);
        const double a = s / c;
        // the product the kernel computes, entry by entry
        const cd L0 = l0 * c, L1 = l1 * c;
        const cd q00 = L0, q01 = -L0 * a * r1, q10 = L1 * a, q11 = L1 * r1;
        const double err = std::max(std::max(std::abs(q00 - m00), std::abs(q01 - m01)),
                                    std::max(std::abs(q10 - m10), std::abs(q11 - m11)));
        if (!(err <= 1e-14)) {
            prog.ops.push_back(op);
            continue;
        }
        if (r1 != cd(1.0, 0.0)) prog.ops.push_back(diag_on(prog, op.sel, op.pc, cd(1.0, 0.0), r1));
        FOp t = op;
        t.type = OP_T2;
        const cd av(a, 0.0);
        t.coef = add_coef(prog, &av, 1);
        prog.ops.push_back(t);
        if (L0 != cd(1.0, 0.0) || L1 != cd(1.0, 0.0)) prog.ops.push_back(diag_on(prog, op.sel, op.pc, L0, L1));
    }
}

// Estimated vector instructions per thread of one op: FP64 FMAs per
// output amplitude times the outputs it writes, plus a fixed cost for the
// op's decode (dependent scalar loads, branches) and for moving its
// results back into the registers the op loop carries.
constexpr uint32_t kOpOverhead = 80;
uint32_t op_cost(const FOp& op) {
    const uint32_t outs = (op.type == OP_D || op.type == OP_S) ? (uint32_t)kRegs : (uint32_t)popc(op.pc & 0xFFFFu);
    uint32_t per = 0;
    switch (op.type) {
    case OP_D: per = has_ext(op) ? 8 : 4; break;  // a table per variant: per-lane loads or a second product
    case OP_M2: per = 8; break;
    case OP_M4: per = 16; break;
    case OP_P4: per = 8; break;
    case OP_P4R: per = 4; break;
    case OP_A2: per = 4; break;
    case OP_X2: per = 1; break;
    case OP_G: per = 0; break;
    case OP_S: per = ((op.sel & 1u) ? 4 : 0) + 4 * 2 + ((op.sel & 2u) ? 4 : 0); break;
    case OP_PH: return 4 * kRegs / 2 + kOpOverhead;
    default: per = 2; break;  // T2
    }
    return per * outs + kOpOverhead;
}

// Picks between a phase with and without split 2x2s.
uint32_t phase_cost(const Program& prog, uint32_t begin) {
    uint32_t cost = 0;
    for (size_t o = begin; o < prog.ops.size(); ++o) cost += op_cost(prog.ops[o]);
    return cost;
}

// Folds runs [plain diagonal] [T2 on distinct register bits, no
// controls] [plain diagonal] into one S op where the estimate says it is
// cheaper (missing bits a = 0; a missing diagonal is an identity table the
// kernel skips, flagged in sel: bit 0 first, bit 1 second).
void fold_structured(Program& prog, uint32_t begin) {
    const std::vector<FOp> in(prog.ops.begin() + begin, prog.ops.end());
    prog.ops.resize(begin);
    const size_t n = in.size();
    size_t i = 0;
    while (i < n) {
        size_t j = i;
        const FOp* dr = nullptr;
        const FOp* dl = nullptr;
        if (plain_diag(in[j])) dr = &in[j++];
        uint32_t bits = 0;
        double t[kRegBits] = {0.0, 0.0, 0.0, 0.0};
        while (j < n && in[j].type == OP_T2 && in[j].pc == 0xFFFFu && !in[j].lctrl && !in[j].gctrl &&
               !((bits >> in[j].sel) & 1u)) {
            bits |= 1u << in[j].sel;
            t[in[j].sel] = prog.coef[in[j].coef];
            ++j;
        }
        if (bits && j < n && plain_diag(in[j])) dl = &in[j++];
        uint32_t separate = 0;
        for (size_t k = i; k < j; ++k) separate += op_cost(in[k]);
        FOp sop{};
        sop.type = OP_S;
        sop.pc = 0xFFFFu;
        sop.sel = (dr ? 1u : 0u) | (dl ? 2u : 0u);
        if (!bits || separate <= op_cost(sop)) {
            prog.ops.push_back(in[i]);
            ++i;
            continue;
        }
        cd block[2 * kRegs + kRegBits];
        for (int k = 0; k < kRegs; ++k) {
            block[k] = dr ? cd(prog.coef[dr->coef + 2 * k], prog.coef[dr->coef + 2 * k + 1]) : cd(1.0, 0.0);
            block[kRegs + kRegBits + k] = dl ? cd(prog.coef[dl->coef + 2 * k], prog.coef[dl->coef + 2 * k + 1]) : cd(1.0, 0.0);
        }
        for (int k = 0; k < kRegBits; ++k) block[kRegs + k] = cd(t[k], 0.0);
        sop.coef = add_coef(prog, block, 2 * kRegs + kRegBits);
        prog.ops.push_back(sop);
        i = j;
    }
}

// Complex coefficients an op reads (fused.hip): a diagonal without
// operands outside the registers reads only variant 0 of its table.
uint32_t coef_complex(const FOp& op) {
    switch (op.type) {
    case OP_D: return has_ext(op) ? 4 * kRegs : kRegs;
    case OP_G: return 4 * kRegs;
    case OP_M2: return 4;
    case OP_M4: return 16;
    case OP_P4: return 8;
    case OP_P4R: return 4;
    case OP_A2: return 2;
    case OP_X2: return 0;
    case OP_S: return 2 * kRegs + kRegBits;
    default: return 0;  // T2, PH: inline in the op
    }
}

// A diagonal on register bits whose table is 1 where register bit j is 0
// and one value g where it is 1 becomes PH (8 products instead of 16).
void phase_on_bit(Program& prog, FOp& op) {
    if (op.type != OP_D || has_ext(op)) return;
    const double* t = prog.coef.data() + op.coef;
    for (uint32_t j = 0; j < (uint32_t)kRegBits; ++j) {
        const int i1 = 1 << j;
        bool ok = !(t[2 * i1] == 1.0 && t[2 * i1 + 1] == 0.0);
        for (int i = 0; i < kRegs && ok; ++i) {
            const double* e = t + 2 * i;
            ok = (i >> j) & 1 ? (e[0] == t[2 * i1] && e[1] == t[2 * i1 + 1]) : (e[0] == 1.0 && e[1] == 0.0);
        }
        if (!ok) continue;
        const cd g(t[2 * i1], t[2 * i1 + 1]);
        op.type = OP_PH;
        op.sel = j;
        op.coef = add_coef(prog, &g, 1);
        return;
    }
}

// Finishes a register phase's ops [begin, end): diagonals that depend on
// bits outside the tile only become G ops at the end (a per-tile scalar
// commutes with every op of the tile); the other diagonals are merged, with
// and without split 2x2s, whichever is cheaper; then the ratio tables.
void finish_phase(Program& prog, uint32_t begin, size_t coef_begin, const FPhase& ph) {
    std::vector<FOp> scalars;
    // Every variant below starts from the same tables and appends the same
    // new ones at the same offsets before merging (tables are rewritten
    // only while merging), so the merges share one cache of their deps.
    DepCache deps;
    {
        std::vector<FOp> rest;
        for (size_t o = begin; o < prog.ops.size(); ++o) {
            const FOp& op = prog.ops[o];
            uint32_t dep = 1;
            if (op.type == OP_D) {
                dep = diag_deps(prog, op);
                deps.push_back({op.coef, dep});
            }
            if (op.type == OP_D && !op.la && !op.lb && !op.lctrl && dep == 0) {
                FOp g = op;
                g.type = OP_G;
                scalars.push_back(g);
            } else {
                rest.push_back(op);
            }
        }
        prog.ops.resize(begin);
        prog.ops.insert(prog.ops.end(), rest.begin(), rest.end());
    }
    // candidates: 2x2s whole or split; diagonals with non-register
    // operands merged into the plain ones or kept apart (so the plain ones
    // can join a structured op)
    // (every table these steps touch or add lies at or after coef_begin)
    const std::vector<FOp> ops0(prog.ops.begin() + begin, prog.ops.end());
    const std::vector<double> coef0(prog.coef.begin() + (long)coef_begin, prog.coef.end());
    std::vector<FOp> best_ops;
    std::vector<double> best_coef;
    uint32_t best = 0xFFFFFFFFu;
    // with every diagonal on register bits only (as the ones splitting adds
    // are), keeping the others apart changes nothing: variant 2 is variant 1
    bool all_plain = true;
    for (const FOp& op : ops0) all_plain = all_plain && (op.type != OP_D || plain_diag(op));
    for (int variant = 0; variant < (all_plain ? 2 : 3); ++variant) {
        prog.ops.resize(begin);
        prog.ops.insert(prog.ops.end(), ops0.begin(), ops0.end());
        prog.coef.resize(coef_begin);
        prog.coef.insert(prog.coef.end(), coef0.begin(), coef0.end());
        if (variant > 0) split_pairs(prog, begin);
        merge_diagonals(prog, begin, variant == 2, deps);
        if (variant > 0) fold_structured(prog, begin);
        const uint32_t cost = phase_cost(prog, begin);
        if (cost < best) {
            best = cost;
            best_ops.assign(prog.ops.begin() + begin, prog.ops.end());
            best_coef.assign(prog.coef.begin() + (long)coef_begin, prog.coef.end());
        }
    }
    prog.ops.resize(begin);
    prog.ops.insert(prog.ops.end(), best_ops.begin(), best_ops.end());
    prog.coef.resize(coef_begin);
    prog.coef.insert(prog.coef.end(), best_coef.begin(), best_coef.end());
    for (size_t o = begin; o < prog.ops.size(); ++o) {
        ratio_table(prog, prog.ops[o], ph);
        phase_on_bit(prog, prog.ops[o]);
    }
    prog.ops.insert(prog.ops.end(), scalars.begin(), scalars.end());
    // T2 and PH carry their one coefficient in the op itself
    for (size_t o = begin; o < prog.ops.size(); ++o) {
        FOp& op = prog.ops[o];
        if (op.type != OP_T2 && op.type != OP_PH) continue;
        const double re = prog.coef[op.coef], im = prog.coef[op.coef + 1];
        std::memcpy(&op.ga, &re, sizeof re);
        std::memcpy(&op.gb, &im, sizeof im);
    }
    // keep only the coefficients the phase's ops read (merged, folded and
    // rewritten tables leave unused blocks behind)
    std::vector<double> kept;
    for (size_t o = begin; o < prog.ops.size(); ++o) {
        FOp& op = prog.ops[o];
        const uint32_t n = 2 * coef_complex(op);
        if (n == 0) { op.coef = 0; continue; }
        const double* src = prog.coef.data() + op.coef;
        op.coef = (uint32_t)(coef_begin + kept.size());
        kept.insert(kept.end(), src, src + n);
    }
    prog.coef.resize(coef_begin);
    prog.coef.insert(prog.coef.end(), kept.begin(), kept.end());
}

struct PhasePlan {
    uint32_t Rl;               // local positions of the phase's register bits
    std::vector<Prim> prims;   // in application order
};

// How a diagonal would be applied in a phase, summed over its operands and
// controls: 0 for a register bit (the table merges with the phase's
// others), 1 for a bit uniform over a wave (outside the tile, or a thread
// bit >= 6), 3 for a bit that differs between the lanes of a wave (each
// lane needs its own table or branch).
int diag_score(const Prim& d, const PhasePlan& pl, const FPass& fp) {
    int score = 0;
    for (uint64_t m = d.bits(); m; m &= m - 1) {
        const int l = local_of(fp.tpos, fp.t, (int)ctz(m));
        if (l < 0) { score += 1; continue; }
        if ((pl.Rl >> l) & 1u) continue;
        const int tbit = l - (int)popc(pl.Rl & ((1u << l) - 1u));
        score += tbit >= 6 ? 1 : 3;
    }
    return score;
}

// Moves each diagonal to the phase of the pass where it is cheapest
// (diag_score), within the phases it can reach: it commutes with every
// primitive except matrix primitives acting on one of its qubits, so it
// may move to the start of any later phase up to the one holding the next
// such primitive, or to the end of any earlier phase back to the one
// holding the previous such primitive.  Ties keep it where it is.
void relocate_diagonals(std::vector<PhasePlan>& plans, const FPass& fp) {
    const int np = (int)plans.size();
    for (int P = 0; P < np; ++P) {
        for (size_t k = 0; k < plans[P].prims.size();) {
            const Prim d = plans[P].prims[k];
            if (d.type != OP_D) { ++k; continue; }
            const uint64_t b = d.bits();
            int hi = np - 1, lo = 0;
            bool found = false;
            for (size_t j = k + 1; j < plans[P].prims.size() && !found; ++j)
                if (plans[P].prims[j].targets() & b) { hi = P; found = true; }
            for (int Q = P + 1; Q < np && !found; ++Q)
                for (const Prim& x : plans[Q].prims)
                    if (x.targets() & b) { hi = Q; found = true; break; }
            found = false;
            for (size_t j = k; j-- > 0 && !found;)
                if (plans[P].prims[j].targets() & b) { lo = P; found = true; }
            for (int Q = P - 1; Q >= 0 && !found; --Q)
                for (const Prim& x : plans[Q].prims)
                    if (x.targets() & b) { lo = Q; found = true; break; }
            int best = P, best_score = diag_score(d, plans[P], fp);
            for (int Q = lo; Q <= hi; ++Q) {
                const int sc = diag_score(d, plans[Q], fp);
                if (sc < best_score || (sc == best_score && Q != best && std::abs(Q - P) < std::abs(best - P))) {
                    best = Q;
                    best_score = sc;
                }
            }
            if (best == P) { ++k; continue; }
            plans[P].prims.erase(plans[P].prims.begin() + (long)k);
            if (best > P) plans[best].prims.insert(plans[best].prims.begin(), d);  // met again when P reaches it
            else plans[best].prims.push_back(d);
        }
    }
}

// Moves the pass's uncontrolled bit flips (x) to its end: a flip of qubit
// q commutes past a later primitive once that primitive's matrix or table
// is conjugated by it (index bit q of its rows and columns inverted), so
// the flips pile up in *flip and the pass applies them as part of its
// final LDS read-out, at no cost.  A primitive controlled by a pending
// flip's qubit would need an anti-control: the flip is applied there.
std::vector<Prim> defer_flips(const std::vector<Prim>& in, uint64_t* flip) {
    std::vector<Prim> out;
    out.reserve(in.size());
    uint64_t F = 0;
    for (const Prim& p0 : in) {
        if (p0.type == OP_M2 && p0.ctrl == 0 && p0.m[0] == 0.0 && p0.m[1] == 1.0 && p0.m[2] == 1.0 && p0.m[3] == 0.0) {
            F ^= 1ull << p0.q0;
            continue;
        }
        for (uint64_t m = p0.ctrl & F; m; m &= m - 1) {
            const int q = (int)ctz(m);
            out.push_back(m2(q, 0.0, 1.0, 1.0, 0.0, 0));
            F &= ~(1ull << q);
        }
        Prim p = p0;
        const int f = (int)((F >> p.q0) & 1u) | (p.q1 >= 0 ? (int)((F >> p.q1) & 1u) << 1 : 0);
        if (f) {
            if (p.type == OP_M2) {
                std::swap(p.m[0], p.m[3]);
                std::swap(p.m[1], p.m[2]);
            } else if (p.type == OP_M4) {
                for (int r = 0; r < 4; ++r)
                    for (int c = 0; c < 4; ++c) p.m[4 * r + c] = p0.m[4 * (r ^ f) + (c ^ f)];
            } else {  // D: table entry k = a | 2b
                for (int k = 0; k < 4; ++k) p.m[k] = p0.m[k ^ f];
            }
        }
        out.push_back(p);
    }
    *flip = F;
    return out;
}

void compile_pass(Program& prog, const PassPlan& pp_in, unsigned dense_min) {
    uint64_t flip = 0;
    PassPlan pp;
    pp.T = pp_in.T;
    pp.permuted = pp_in.permuted;
    std::memcpy(pp.sigma, pp_in.sigma, sizeof pp.sigma);
    pp.prims = defer_flips(pp_in.prims, &flip);
    FPass fp{};
    fp.t = popc(pp.T);
    fp.nthreads = 1u << (fp.t - kRegBits);
    {
        unsigned k = 0;
        for (uint64_t m = pp.T; m; m &= m - 1) fp.tpos[k++] = ctz(m);
        for (k = 0; k < fp.t; ++k) fp.tpack[k / 4] |= fp.tpos[k] << (8 * (k % 4));
    }
    for (int j = 0; j < kRegs; ++j) {
        uint64_t off = 0;
        for (int k = 0; k < kRegBits; ++k)
            if (j >> k & 1) off |= 1ull << fp.tpos[fp.t - kRegBits + k];
        fp.hoff[j] = off;
    }
    const uint32_t top_regs = ((1u << kRegBits) - 1u) << (fp.t - kRegBits);  // local bits t-4..t-1

    // phases: greedy register sets of four local qubits
    // (the queue holds indices into pp.prims; a scan stops once the
    // left-behind matrix primitives target every qubit the pass touches)
    const std::vector<Prim>& all = pp.prims;
    uint64_t support = 0;
    for (const Prim& p : all) support |= p.bits();
    std::vector<uint32_t> rest(all.size()), leave;
    for (size_t k = 0; k < all.size(); ++k) rest[k] = (uint32_t)k;
    std::vector<PhasePlan> plans;
    while (!rest.empty()) {
        uint32_t Rl = 0;  // local register bits
        Left left;
        PhasePlan plan;
        plan.prims.reserve(rest.size());
        leave.clear();
        size_t i = 0;
        for (; i < rest.size() && (left.mat_targets & support) != support; ++i) {
            const Prim& p = all[rest[i]];
            if (left.blocks(p)) { left.add(p); leave.push_back(rest[i]); continue; }
            if (p.type == OP_D) { plan.prims.push_back(p); continue; }
            uint32_t need = 0;
            for (uint64_t m = p.targets(); m; m &= m - 1) need |= 1u << local_of(fp.tpos, fp.t, (int)ctz(m));
            if (popc(Rl | need) <= kRegBits) { Rl |= need; plan.prims.push_back(p); }
            else { left.add(p); leave.push_back(rest[i]); }
        }
        leave.insert(leave.end(), rest.begin() + (long)i, rest.end());
        // complete the register set, preferring the top local bits (direct HBM access)
        for (int k = (int)fp.t - 1; k >= 0 && popc(Rl) < kRegBits; --k)
            if ((top_regs >> k) & 1) Rl |= 1u << k;
        for (int k = 0; k < (int)fp.t && popc(Rl) < kRegBits; ++k) Rl |= 1u << k;
        plan.Rl = Rl;
        plans.push_back(std::move(plan));
        rest.swap(leave);
    }
    if (plans.empty()) {  // a pass that only permutes: one phase without ops
        PhasePlan plan;
        plan.Rl = top_regs;
        plans.push_back(std::move(plan));
    }
    relocate_diagonals(plans, fp);

    fp.phase_begin = (uint32_t)prog.phases.size();
    for (PhasePlan& plan : plans) {
        if (plan.prims.empty() && plans.size() > 1) continue;  // its diagonals all moved
        const uint32_t Rl = plan.Rl;
        const std::vector<Prim>& take = plan.prims;
        FPhase ph = make_phase(Rl);
        auto regbit = [&](int q) -> int {
            const int l = local_of(fp.tpos, fp.t, q);
            if (l < 0) return -1;
            for (int k = 0; k < kRegBits; ++k)
                if ((int)ph.rpos[k] == l) return k;
            return -1;
        };
        ph.op_begin = (uint32_t)prog.ops.size();
        const size_t coef_begin = prog.coef.size();
        bool foldable = dense_min > 0 && fp.t >= (uint32_t)kMinDenseTile && take.size() >= dense_min;
        for (const Prim& p0 : take) {
            Prim p = p0;
            // a controlled phase diag(1, g) on q0 is symmetric in q0 and its
            // controls: make a register bit the operand where one is
            if (p.type == OP_D && p.q1 < 0 && p.m[0] == cd(1.0, 0.0) && p.m[2] == cd(1.0, 0.0) && regbit(p.q0) < 0) {
                for (uint64_t m = p.ctrl; m; m &= m - 1) {
                    const int c = (int)ctz(m);
                    if (regbit(c) >= 0) {
                        p.ctrl = (p.ctrl & ~(1ull << c)) | (1ull << p.q0);
                        p.q0 = c;
                        break;
                    }
                }
            }
            FOp op{};
            op.type = p.type;
            // controls: register / lane / outside-tile
            uint32_t rctrl = 0;  // register bits that must be 1
            for (uint64_t m = p.ctrl; m; m &= m - 1) {
                const int q = (int)ctz(m);
                const int j = regbit(q);
                const int l = local_of(fp.tpos, fp.t, q);
                if (j >= 0) rctrl |= 1u << j;
                else if (l >= 0) op.lctrl |= 1u << l;
                else op.gctrl |= 1ull << q;
            }
            uint32_t pc = 0;
            for (int i = 0; i < kRegs; ++i)
                if (((uint32_t)i & rctrl) == rctrl) pc |= 1u << i;
            op.pc = pc;
            if (p.type == OP_D) {
                // operand a / b: register bit, lane bit, bit outside the tile, or absent
                int ja = -1, jb = -1;
                auto operand = [&](int q, int& j, uint32_t& lm, uint64_t& gm) {
                    if (q < 0) return;
                    j = regbit(q);
                    if (j >= 0) return;
                    const int l = local_of(fp.tpos, fp.t, q);
                    if (l >= 0) lm = 1u << l;
                    else gm = 1ull << q;
                };
                operand(p.q0, ja, op.la, op.ga);
                operand(p.q1, jb, op.lb, op.gb);
                cd tab[4 * kRegs];
                for (int k = 0; k < 4; ++k)
                    for (int i = 0; i < kRegs; ++i) {
                        const int A = ja >= 0 ? (i >> ja) & 1 : (k & 1);
                        const int B = jb >= 0 ? (i >> jb) & 1 : (k >> 1);
                        tab[k * kRegs + i] = ((pc >> i) & 1u) ? p.m[A | (B << 1)] : cd(1.0, 0.0);
                    }
                op.coef = add_coef(prog, tab, 4 * kRegs);
            } else if (p.type == OP_M2) {
                op.sel = (uint32_t)regbit(p.q0);
                if (p.m[0] == 0.0 && p.m[3] == 0.0) {
                    if (p.m[1] == 1.0 && p.m[2] == 1.0) {
                        op.type = OP_X2;
                        op.coef = 0;
                    } else {
                        op.type = OP_A2;
                        const cd a[2] = {p.m[1], p.m[2]};
                        op.coef = add_coef(prog, a, 2);
                    }
                } else {
                    op.coef = add_coef(prog, p.m, 4);
                }
            } else {
                int j0 = regbit(p.q0), j1 = regbit(p.q1);
                cd u[16];
                if (j0 < j1) {
                    for (int k = 0; k < 16; ++k) u[k] = p.m[k];
                } else {  // swap the roles of matrix bits 0 and 1
                    auto sw = [](int x) { return ((x & 1) << 1) | ((x >> 1) & 1); };
                    for (int r = 0; r < 4; ++r)
                        for (int c = 0; c < 4; ++c) u[4 * sw(r) + sw(c)] = p.m[4 * r + c];
                    std::swap(j0, j1);
                }
                op.sel = (uint32_t)(4 * j0 + j1);
                bool paired = true;  // u[r][c] == 0 unless c is r or r ^ 3
                for (int r = 0; r < 4; ++r)
                    for (int c = 0; c < 4; ++c)
                        if (c != r && c != (r ^ 3) && u[4 * r + c] != 0.0) paired = false;
                bool real_imag = paired;  // real diagonal, imaginary pairing (rxx, ryy)
                for (int r = 0; r < 4 && real_imag; ++r)
                    real_imag = u[4 * r + r].imag() == 0.0 && u[4 * r + (r ^ 3)].real() == 0.0;
                if (real_imag) {
                    op.type = OP_P4R;
                    cd pr[4];
                    for (int r = 0; r < 4; ++r) pr[r] = cd(u[4 * r + r].real(), u[4 * r + (r ^ 3)].imag());
                    op.coef = add_coef(prog, pr, 4);
                } else if (paired) {
                    op.type = OP_P4;
                    cd pr[8];
                    for (int r = 0; r < 4; ++r) {
                        pr[2 * r] = u[4 * r + r];
                        pr[2 * r + 1] = u[4 * r + (r ^ 3)];
                    }
                    op.coef = add_coef(prog, pr, 8);
                } else {
                    op.coef = add_coef(prog, u, 16);
                }
            }
            if (op.lctrl || op.gctrl || op.la || op.lb || op.ga || op.gb) foldable = false;
            prog.ops.push_back(op);
        }
        if (!foldable) finish_phase(prog, ph.op_begin, coef_begin, ph);
        ph.op_end = (uint32_t)prog.ops.size();
        if (foldable) {
            cd U[kRegs * kRegs];
            dense_of(prog, ph.op_begin, ph.op_end, U);
            prog.ops.resize(ph.op_begin);
            prog.coef.resize(coef_begin);
            ph.op_end = ph.op_begin;
            ph.dense = 1;
            ph.ucoef = add_coef(prog, U, kRegs * kRegs);
        }
        prog.phases.push_back(ph);
    }
    fp.phase_end = (uint32_t)prog.phases.size();
    auto is_top = [&](const FPhase& ph) {
        if (ph.dense) return false;  // dense phases work from the LDS tile
        for (int k = 0; k < kRegBits; ++k)
            if (ph.rpos[k] != fp.t - kRegBits + k) return false;
        return true;
    };
    fp.direct_first = is_top(prog.phases[fp.phase_begin]) ? 1u : 0u;
    fp.direct_last = is_top(prog.phases[fp.phase_end - 1]) ? 1u : 0u;
    if (pp.permuted || flip) {
        // the deferred flips and the qubit permutation as a permuted LDS
        // read-out of the tile: the amplitude at local index e belongs at
        // lam(e ^ f) (f: the flipped local bits; lam moves local bit l,
        // index bit tpos[l], to the local bit of index bit sigma[tpos[l]]),
        // so the thread that stores local index L reads local index
        // lam^-1(L) ^ f, whose swizzled offset is XOR-linear in L's bits
        // plus the constant swizzle of f
        int inv[kMaxTile];
        for (uint32_t l = 0; l < fp.t; ++l) inv[local_of(fp.tpos, fp.t, pp.sigma[fp.tpos[l]])] = (int)l;
        uint32_t f = 0;
        for (uint64_t m = flip; m; m &= m - 1) f |= 1u << local_of(fp.tpos, fp.t, (int)ctz(m));
        fp.permuted = 1;
        fp.direct_last = 0;
        for (uint32_t k = 0; k + kRegBits < fp.t; ++k) fp.pimg[k] = lds_swizzle(1u << inv[k]);
        for (int j = 0; j < kRegs; ++j) {
            uint32_t e = 0;
            for (int b = 0; b < kRegBits; ++b)
                if (j >> b & 1) e |= 1u << inv[fp.t - kRegBits + b];
            fp.pout[j] = lds_swizzle(e ^ f);
        }
    }
    prog.passes.push_back(fp);
}

// Appends `src` (offsets from 0) to `dst`, moving its phase, op and
// coefficient offsets past what `dst` holds.
void append_program(Program& dst, const Program& src) {
    const uint32_t pb = (uint32_t)dst.phases.size(), ob = (uint32_t)dst.ops.size(), cb = (uint32_t)dst.coef.size();
    for (FPass fp : src.passes) {
        fp.phase_begin += pb;
        fp.phase_end += pb;
        dst.passes.push_back(fp);
    }
    for (FPhase ph : src.phases) {
        ph.op_begin += ob;
        ph.op_end += ob;
        if (ph.dense) ph.ucoef += cb;
        dst.phases.push_back(ph);
    }
    for (FOp op : src.ops) {
        if (coef_complex(op)) op.coef += cb;
        dst.ops.push_back(op);
    }
    dst.coef.insert(dst.coef.end(), src.coef.begin(), src.coef.end());
}

// With parallel, passes are compiled on up to kCompileThreads threads (the
// caller's among them) when the queue holds at least kParallelPrims
// primitives to compile.
constexpr unsigned kCompileThreads = 8;
constexpr size_t kParallelPrims = 64;

template <class F>
void parallel_for(size_t n, bool parallel, const F& fn) {
    const unsigned hw = std::thread::hardware_concurrency();
    size_t nt = parallel ? std::min<size_t>(n, std::min(hw ? hw : 1u, kCompileThreads)) : 1;
    if (nt <= 1) {
        for (size_t k = 0; k < n; ++k) fn(k);
        return;
    }
    std::atomic<size_t> next{0};
    auto run = [&] {
        for (size_t k; (k = next.fetch_add(1, std::memory_order_relaxed)) < n;) fn(k);
    };
    std::vector<std::thread> pool;
    pool.reserve(nt - 1);
    for (size_t t = 1; t < nt; ++t) pool.emplace_back(run);
    run();
    for (std::thread& t : pool) t.join();
}

// Fitted run-time model of a compiled queue at 2^30 amplitudes (seconds per
// pass, seconds per unit of program_cost), per tile size 10, 11, 12, 13:
// least squares over four measured workloads (30-qubit brickwork per layer
// and whole, 30-qubit TFIM ring per step and whole, 50 steps), each within
// 3 %.  FP64 has no tile of 13 (its entry is never read).
struct TileModel {
    double pass_s, cost_s;
};
constexpr TileModel kModelF64[] = {{5.320e-3, 1.217e-6}, {5.298e-3, 1.258e-6}, {5.396e-3, 1.547e-6}, {0.0, 0.0}};
constexpr TileModel kModelF32[] = {{2.881e-3, 6.410e-7}, {2.401e-3, 9.295e-7}, {2.505e-3, 9.065e-7},
                                   {3.068e-3, 8.457e-7}};
static_assert(sizeof(kModelF64) / sizeof(kModelF64[0]) == sizeof(kAutoTiles) / sizeof(kAutoTiles[0]), "model");
static_assert(sizeof(kModelF32) / sizeof(kModelF32[0]) == sizeof(kAutoTiles) / sizeof(kAutoTiles[0]), "model");

}  // namespace

uint64_t program_cost(const Program& prog) {
    uint64_t cost = 0;
    for (const FOp& op : prog.ops) cost += op_cost(op);
    return cost;
}

unsigned plan_auto(const qvnt_gate* g, size_t n, unsigned ns, bool fp64, std::vector<Step>& steps, bool single_plain,
                   bool parallel) {
    constexpr size_t kTiles = sizeof(kAutoTiles) / sizeof(kAutoTiles[0]);
    std::vector<Step> cand[kTiles];
    double est[kTiles];
    // the candidates compile side by side
    parallel_for(kTiles, true, [&](size_t k) {
        est[k] = -1.0;
        if (kAutoTiles[k] > ns || kAutoTiles[k] > max_tile(fp64)) return;
        plan(g, n, ns, kAutoTiles[k], cand[k], 0, single_plain, parallel, fp64 ? 16 : 8);
        const TileModel& m = (fp64 ? kModelF64 : kModelF32)[k];
        double passes = 0.0, cost = 0.0;
        for (const Step& st : cand[k]) {
            if (st.plain) {
                passes += 1.0;
                continue;
            }
            passes += (double)st.prog.passes.size();
            cost += (double)program_cost(st.prog);
        }
        est[k] = passes * m.pass_s + cost * m.cost_s;
    });
    size_t best = kTiles;
    for (size_t k = 0; k < kTiles; ++k)
        if (est[k] >= 0.0 && (best == kTiles || est[k] < est[best])) best = k;
    if (best == kTiles) {  // a buffer below every candidate tile
        plan(g, n, ns, ns, steps, 0, single_plain, parallel, fp64 ? 16 : 8);
        return ns;
    }
    steps.swap(cand[best]);
    return kAutoTiles[best];
}

namespace {

// Takes the swap marks out of a segment's primitives: an uncontrolled swap
// of qubits a and b exchanges their amplitude bits, so instead of moving
// amplitudes the compiler exchanges the two qubits' positions for every
// later primitive.  Afterwards qubit q's bit sits at position pos[q];
// *want (indexed by position) says where each position's bit belongs.
// Returns false (want untouched) if the segment has no swap.  With
// `keep`, the marks become swap matrices again instead (tiles too small
// for the permutation passes).
bool relabel_swaps(std::vector<Prim>& prims, unsigned ns, bool keep, std::vector<int>* want) {
    bool any = false;
    for (const Prim& p : prims) any |= p.type == kRelabel;
    if (!any) return false;
    if (keep) {
        for (Prim& p : prims) {
            if (p.type != kRelabel) continue;
            cd u[16] = {};
            u[0] = 1.0; u[4 * 1 + 2] = 1.0; u[4 * 2 + 1] = 1.0; u[15] = 1.0;
            const long src = p.src;
            p = m4(p.q0, p.q1, u, 0);
            p.src = src;
        }
        return false;
    }
    int pos[64];
    for (int q = 0; q < 64; ++q) pos[q] = q;
    std::vector<Prim> out;
    out.reserve(prims.size());
    for (Prim p : prims) {
        if (p.type == kRelabel) {
            std::swap(pos[p.q0], pos[p.q1]);
            continue;
        }
        remap(p, pos);
        out.push_back(p);
    }
    prims.swap(out);
    want->assign(ns, 0);
    for (unsigned q = 0; q < ns; ++q) (*want)[(size_t)pos[q]] = (int)q;
    return true;
}

}  // namespace

void plan(const qvnt_gate* g, size_t n, unsigned ns, unsigned tq, std::vector<Step>& steps, unsigned dense_min,
          bool single_plain, bool parallel, unsigned amp_bytes) {
    const uint64_t lmask = (1ull << ns) - 1u;
    size_t i = 0;
    while (i < n) {
        // a segment of fusable gates, then (if any) one gate for the plain kernels
        std::vector<Prim> prims;
        size_t j = i;
        for (; j < n; ++j) {
            if ((g[j].ctrl & ~lmask) || g[j].kind == QVNT_ID) continue;
            const size_t before = prims.size();
            if (!lower(g[j], lmask, prims)) break;
            for (size_t k = before; k < prims.size(); ++k) prims[k].src = (long)j;
        }
        if (!prims.empty()) {
            const unsigned rows = row_bits(amp_bytes);
            std::vector<int> want;
            const bool relabeled = relabel_swaps(prims, ns, ns > tq && tq < rows + 2, &want);
            const std::vector<Prim> merged = merge_single(prims);
            std::vector<PassPlan> passes = schedule_passes(merged, ns, tq, rows, relabeled ? &want : nullptr);
            // the backward plan where it needs fewer passes
            if (!relabeled && passes.size() > 1) {
                std::vector<PassPlan> back = schedule_backward(merged, ns, tq, rows);
                if (back.size() < passes.size()) passes.swap(back);
            }
            // primitives per source SingleOp, to find passes that hold one
            // whole SingleOp and nothing else: one plain launch moves at
            // most the same bytes as a pass (fewer with controls) and skips
            // the tile's LDS round trips
            std::vector<uint32_t> per_src(n, 0);
            for (const PassPlan& pp : passes)
                for (const Prim& p : pp.prims)
                    if (p.src >= 0) ++per_src[(size_t)p.src];
            std::vector<long> single_src(passes.size(), -1);
            size_t work = 0;
            for (size_t k = 0; k < passes.size(); ++k) {
                const PassPlan& pp = passes[k];
                const long src = pp.prims.empty() ? -1 : pp.prims[0].src;
                // (not in a relabeled segment: its primitives address
                // moved qubits, the SingleOp its original ones)
                bool single = single_plain && !relabeled && src >= 0 && per_src[(size_t)src] == pp.prims.size();
                for (const Prim& p : pp.prims) single = single && p.src == src;
                if (single) single_src[k] = src;
                else work += pp.prims.size();
            }
            // passes compile independently, each into a program of its own
            // (offsets from 0), concatenated in order afterwards
            std::vector<Program> progs(passes.size());
            parallel_for(passes.size(), parallel && work >= kParallelPrims, [&](size_t k) {
                if (single_src[k] < 0) compile_pass(progs[k], passes[k], dense_min);
            });
            Step st;
            for (size_t k = 0; k < passes.size(); ++k) {
                if (single_src[k] < 0) {
                    append_program(st.prog, progs[k]);
                    continue;
                }
                if (!st.prog.passes.empty()) steps.push_back(std::move(st));
                st = Step();
                Step one;
                one.plain = true;
                one.gate = g[single_src[k]];
                steps.push_back(std::move(one));
            }
            if (!st.prog.passes.empty()) steps.push_back(std::move(st));
        }
        if (j < n) {
            Step st;
            st.plain = true;
            st.gate = g[j];
            steps.push_back(std::move(st));
            ++j;
        }
        i = j;
    }
}

}  // namespace fusion
}  // namespace qvnt
