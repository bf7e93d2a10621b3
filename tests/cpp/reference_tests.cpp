// The reference's own unit tests and doctests, restated against the C++
// facade (include/qvnt.hpp) over the device library.  Mode "host" runs the
// operator tests (no GPU needed); mode "device" runs the register tests.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "qvnt.hpp"

using namespace qvnt;
static int failures = 0;
#define EXPECT(c) do { if (!(c)) { std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); ++failures; } } while (0)

// src/operator/single/mod.rs:142-152 (name)
static void single_name() {
    MultiOp o = op::x(123);
    EXPECT(o.name() == "[X123]");
    EXPECT(o.c(4)->name() == "[C4_X123]");
}
// src/operator/single/mod.rs:154-161 (wrong_ctrl_mask)
static void wrong_ctrl_mask() {
    MultiOp o = op::ryy(1.35, 0b101);
    EXPECT((o.act_on() & 0b001) != 0);
    EXPECT(!o.c(0b001).has_value());
    EXPECT((o.act_on() & 0b010) == 0);
    EXPECT(o.c(0b010).has_value());
}
// src/operator/multi/mod.rs:201-208 (ops)
static void multi_ops() {
    MultiOp pend = op::id() * *op::h(0b001).c(0b010) * *op::x(0b011).c(0b100) * op::rz(5.0, 0b001);
    EXPECT(pend.len() == 3);
}
// README.md:95-98 (should_panic: overlapping control)
static void readme_overlap() { EXPECT(!op::x(0b001).c(0b001).has_value()); }
// src/operator/mod.rs:151-156 (expect on a wrong bit count)
static void rotation_mask_panics() {
    bool threw = false;
    try { op::rx(1.0, 0b11); } catch (const Error& e) { threw = e.code == QVNT_E_INVALID_MASK; }
    EXPECT(threw);
}
// src/register/class.rs:14-31 (doctests: with_state, tensor product) and
// :37-48 (get, Debug); class.rs:159-165 (display)
static void creg_doc() {
    CReg c0 = CReg::with_state(4, 11);
    CReg c1 = CReg::with_state(4, 7);
    CReg c = c0 * c1;
    EXPECT(c == CReg::with_state(8, 123));
    EXPECT(c.get() == 123);
    EXPECT(c.debug() == "(01111011)");
    EXPECT(CReg::with_state(17, 123).debug() == "(00000000001111011)");
    CReg d = CReg::with_state(4, 11);
    d *= CReg::with_state(4, 7);
    EXPECT(d == c);
}
// src/register/class.rs:85-125: reset, set, xor, get_by_mask
static void creg_bits() {
    CReg c = CReg::new_(6);
    c.reset(0b1101011);                    // masked to 6 bits
    EXPECT(c.get() == 0b101011);
    c.set(true, 0b010100);
    EXPECT(c.get() == 0b111111);
    c.set(false, 0b000011);
    EXPECT(c.get() == 0b111100);
    c.xor_(true, 0b100001);
    EXPECT(c.get() == 0b011101);
    c.xor_(false, 0b111111);               // bit false: unchanged
    EXPECT(c.get() == 0b011101);
    // bits 0, 2, 3 of 0b011101 are 1, 1, 1; bit 5 (mask beyond 6 bits drops bit 7) is 0
    EXPECT(c.get_by_mask(0b10101101) == 0b0111);
    EXPECT(c.get_by_mask(0) == 0);
    c.set_num(3);
    EXPECT(c.num() == 3 && c.get() == 0b011101);   // set_num keeps the value
    EXPECT(c.get_by_mask(0b111) == 0b101);
}
// src/register/virtl.rs:161-178 (index)
static void vreg_index() {
    auto even = [](size_t i) { return i % 2 == 0; };
    auto odd = [](size_t i) { return i % 2 != 0; };
    VReg v = VReg::new_(8);
    EXPECT(v[3] == 0b00001000);
    EXPECT(v[even] == 0b01010101);
    EXPECT(v[odd] == 0b10101010);
    EXPECT(v.all() == 0b11111111);
    EXPECT(v[([](size_t) { return true; })] == 0b11111111);
    EXPECT((v[{0, 7}]) == 0b10000001);
}
// src/register/virtl.rs:14-84 (doctests)
static void vreg_doc() {
    VReg v = VReg::new_(8);
    EXPECT((v[{3, 4, 5}]) == 0b00111000);
    EXPECT(v[([](size_t i) { return i % 2 != 0; })] == 0b10101010);
    VReg e = VReg::from(v[([](size_t i) { return i % 2 == 0; })]);
    VReg o = VReg::from(v[([](size_t i) { return i % 2 != 0; })]);
    EXPECT(e.all() == 0b01010101 && o.all() == 0b10101010);
    EXPECT(e.len() == 4 && e[1] == 0b100);
    EXPECT(VReg::from(0b101).debug() == "[00000000000000000000000000000001, 00000000000000000000000000000004]");
}
// src/register/quant.rs:642-677 (quantum_reg)
static void quantum_reg() {
    QReg reg = QReg::with_state(4, 0b1100);
    const uint64_t mask = 0b0110;
    MultiOp operator_ = op::h(0b1111) * *op::h(0b0011).c(0b1000) * *op::swap(0b1001).c(0b0010);
    reg.apply(operator_);
    EXPECT(operator_.name() == "[H3, H12, C8_H3, C2_SWAP9]");
    const double want[16] = {0.25, 0.25, 0.25, 0.0, -0.25, -0.25, -0.25, 0.0, -0.5, 0.0, 0.25, 0.0, 0.5, 0.0, -0.25, 0.0};
    auto psi = reg.get_state();
    for (int i = 0; i < 16; ++i) EXPECT(psi[i] == std::complex<double>(want[i], 0.0));
    EXPECT(reg.debug() ==
           "QReg { 0: Complex { re: 0.25, im: 0.0 }, 1: Complex { re: 0.25, im: 0.0 }, 2: Complex { re: 0.25, im: 0.0 }, "
           "3: Complex { re: 0.0, im: 0.0 }, 4: Complex { re: -0.25, im: 0.0 }, 5: Complex { re: -0.25, im: 0.0 }, "
           "6: Complex { re: -0.25, im: 0.0 }, 7: Complex { re: 0.0, im: 0.0 }, .. }");
    EXPECT((reg.measure_mask(mask).get() & ~mask) == 0);
}
// src/register/quant.rs:81-96 (doctest: Bell state)
static void bell() {
    QReg q = QReg::new_(2);
    MultiOp gate = op::h(0b01) * *op::x(0b10).c(0b01);
    q.apply(gate);
    auto p = q.get_probabilities();
    EXPECT(p.size() == 4 && p[0] == 0.5 && p[1] == 0.0 && p[2] == 0.0 && p[3] == 0.5);
}
// src/register/quant.rs:679-711 (tensor, the reg3 part)
static void tensor_reg3() {
    QReg reg3 = QReg::with_state(3, 0b101);
    reg3.apply(op::h(0b101));
    auto p = reg3.get_probabilities();
    const double want[8] = {0.25, 0.25, 0, 0, 0.25, 0.25, 0, 0};
    for (int i = 0; i < 8; ++i) EXPECT(std::fabs(p[i] - want[i]) < 1e-9);
}
// src/register/quant.rs:679-711 (tensor, the reg1 * reg2 part)
static void tensor_product() {
    QReg reg1 = QReg::with_state(2, 0b01);
    QReg reg2 = QReg::with_state(1, 0b1);
    reg1.apply(op::h(0b01));
    reg2.apply(op::h(0b01));
    auto p = (reg1 * reg2).get_probabilities();
    const double want[8] = {0.25, 0.25, 0, 0, 0.25, 0.25, 0, 0};
    EXPECT(p.size() == 8);
    for (int i = 0; i < 8; ++i) EXPECT(std::fabs(p[i] - want[i]) < 1e-9);
}
// src/register/quant.rs:713-724 (histogram)
static void histogram() {
    QReg q = QReg::with_state(8, 123);
    q.apply(op::h(255));
    for (uint64_t k = 0; k < 10; ++k) {
        auto hist = q.sample_all(2048, k);
        uint64_t sum = 0;
        for (uint64_t c : hist) sum += c;
        EXPECT(hist.size() == 256 && sum == 2048);
    }
}
// src/register/virtl.rs:180-218 (virtual_regs): reg = new(0) * new(3) * new(2)
static void virtual_regs_sizes() {
    QReg a = QReg::new_(0), b = QReg::new_(3), c = QReg::new_(2);
    QReg ab = a * b;
    QReg reg = ab * c;
    EXPECT(reg.num() == 5 && reg.get_probabilities()[0] == 1.0);
    VReg r = reg.get_vreg();
    std::optional<VReg> m = reg.get_vreg_by(0b01101), x = reg.get_vreg_by(0b00111), y = reg.get_vreg_by(0b11000);
    EXPECT(m && x && y && !reg.get_vreg_by(0b100000));
    EXPECT(r[0] == 0b00001 && r[1] == 0b00010 && r[2] == 0b00100 && r[3] == 0b01000 && r[4] == 0b10000);
    EXPECT(r.all() == 0b11111);
    EXPECT((*m)[0] == 0b00001 && (*m)[1] == 0b00100 && (*m)[2] == 0b01000 && m->all() == 0b01101);
    EXPECT((*x)[0] == 0b00001 && (*x)[1] == 0b00010 && (*x)[2] == 0b00100 && x->all() == 0b00111);
    EXPECT((*y)[0] == 0b01000 && (*y)[1] == 0b10000 && y->all() == 0b11000);
}
// quant.rs:603-623: a register of at most 8 amplitudes prints them all, without ".."
static void debug_small() {
    QReg q = QReg::with_state(1, 1);
    EXPECT(q.debug() == "QReg { 0: Complex { re: 0.0, im: 0.0 }, 1: Complex { re: 1.0, im: 0.0 } }");
    QReg z = QReg::new_(0);
    EXPECT(z.debug() == "QReg { 0: Complex { re: 1.0, im: 0.0 } }");
}
// benches/performance.rs:4-11 (perf_test_single) at 18..20 qubits, plain and fused
static void bench_workload() {
    for (uint32_t q = 18; q <= 20; ++q) {
        QReg reg = QReg::with_state(q, 0);
        reg.apply(op::qft(0b0111) * op::qft(0b1110));
        const uint64_t mask = 0b100;
        EXPECT((reg.measure_mask(mask).get() & ~mask) == 0);
        QReg f = QReg::with_state(q, 0);
        f.apply_fused(op::qft(0b0111) * op::qft(0b1110));
        EXPECT(std::fabs(f.get_absolute() - 1.0) < 1e-12);
    }
}

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "host";
    single_name();
    wrong_ctrl_mask();
    multi_ops();
    readme_overlap();
    rotation_mask_panics();
    creg_doc();
    creg_bits();
    vreg_index();
    vreg_doc();
    if (mode == "device") {
        debug_small();
        quantum_reg();
        bell();
        tensor_reg3();
        tensor_product();
        histogram();
        virtual_regs_sizes();
        bench_workload();
        qvnt::cache_trim();  // freed registers' blocks and streams go back
        bell();              // and new registers allocate afresh
    }
    std::printf("%s: %d failure(s)\n", mode.c_str(), failures);
    return failures ? 1 : 0;
}
