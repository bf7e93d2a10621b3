// qvnt.hpp -- C++17 facade over the C ABI (qvnt.h) with the shape of
// qvnt 0.4.4's Rust API:
//   qvnt::op::{id,x,y,z,s,t,rx,ry,rz,rxx,ryy,rzz,phi,swap,sqrt_swap,i_swap,
//              sqrt_i_swap,h,u1,u2,u3,qft,qft_swapped}   src/operator/mod.rs:113-522
//   MultiOp: `a * b` appends, `.c(mask)` -> std::optional (the reference's
//            Option), `.dgr()`, `.act_on()`, Debug name    src/operator/multi/mod.rs
//   QReg::{new_, with_state, apply, apply_fused, get_probabilities,
//          get_absolute, measure_mask, measure}            src/register/quant.rs
//   CReg (classical register, measurement result)           src/register/class.rs
//   VReg (mask building by position / predicate)            src/register/virtl.rs
// Where the reference panics (a rotation mask with the wrong bit count, a
// measurement of a zero state) the facade throws qvnt::Error.
#pragma once

#include <algorithm>
#include <array>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <initializer_list>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "qvnt.h"

namespace qvnt {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error("qvnt error " + std::to_string(c) + ": " + m), code(c) {}
};

inline void check(int rc) {
    if (rc != QVNT_OK) throw Error(rc, qvnt_last_error());
}

// Gives the device blocks and idle streams that freed registers left in the
// library's per-device cache back to the device (qvnt_cache_trim).
inline void cache_trim() { check(qvnt_cache_trim()); }

class MultiOp {
public:
    explicit MultiOp(qvnt_op* h) : h_(h) {}
    MultiOp(const MultiOp& o) : h_(nullptr) { check(qvnt_op_clone(o.h_, &h_)); }
    MultiOp(MultiOp&& o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
    MultiOp& operator=(MultiOp o) noexcept {
        std::swap(h_, o.h_);
        return *this;
    }
    ~MultiOp() { qvnt_op_free(h_); }

    // `*`: rhs runs after *this (MulAssign appends, multi/mod.rs:166-170)
    MultiOp& operator*=(const MultiOp& rhs) {
        check(qvnt_op_mul(h_, rhs.h_));
        return *this;
    }
    friend MultiOp operator*(MultiOp lhs, const MultiOp& rhs) { return lhs *= rhs; }

    // Applicable::c: nullopt where the reference returns None
    std::optional<MultiOp> c(uint64_t mask) const {
        MultiOp out(*this);
        const int rc = qvnt_op_c(out.h_, mask);
        if (rc == QVNT_E_CTRL_OVERLAP) return std::nullopt;
        check(rc);
        return out;
    }
    MultiOp dgr() const {
        MultiOp out(*this);
        check(qvnt_op_dgr(out.h_));
        return out;
    }
    uint64_t act_on() const { return qvnt_op_act_on(h_); }
    size_t len() const { return qvnt_op_len(h_); }
    std::vector<qvnt_gate> gates() const {
        const qvnt_gate* g = qvnt_op_gates(h_);
        return g ? std::vector<qvnt_gate>(g, g + len()) : std::vector<qvnt_gate>();
    }
    std::string name() const {  // the Debug string, e.g. "[H3, H12, C8_H3, C2_SWAP9]"
        size_t need = 0;
        check(qvnt_op_name(h_, nullptr, 0, &need));
        std::string s(need, '\0');
        check(qvnt_op_name(h_, s.data(), need, &need));
        s.resize(need - 1);
        return s;
    }
    const qvnt_op* handle() const { return h_; }

private:
    qvnt_op* h_;
};

namespace op {
namespace detail {
template <class F, class... A>
MultiOp make(F f, A... a) {
    qvnt_op* h = nullptr;
    check(f(a..., &h));
    return MultiOp(h);
}
}  // namespace detail
inline MultiOp id() { return detail::make(qvnt_op_id); }
inline MultiOp x(uint64_t m) { return detail::make(qvnt_op_x, m); }
inline MultiOp y(uint64_t m) { return detail::make(qvnt_op_y, m); }
inline MultiOp z(uint64_t m) { return detail::make(qvnt_op_z, m); }
inline MultiOp s(uint64_t m) { return detail::make(qvnt_op_s, m); }
inline MultiOp t(uint64_t m) { return detail::make(qvnt_op_t, m); }
inline MultiOp rx(double p, uint64_t m) { return detail::make(qvnt_op_rx, p, m); }
inline MultiOp ry(double p, uint64_t m) { return detail::make(qvnt_op_ry, p, m); }
inline MultiOp rz(double p, uint64_t m) { return detail::make(qvnt_op_rz, p, m); }
inline MultiOp rxx(double p, uint64_t m) { return detail::make(qvnt_op_rxx, p, m); }
inline MultiOp ryy(double p, uint64_t m) { return detail::make(qvnt_op_ryy, p, m); }
inline MultiOp rzz(double p, uint64_t m) { return detail::make(qvnt_op_rzz, p, m); }
inline MultiOp phi(double p, uint64_t m) { return detail::make(qvnt_op_phi, p, m); }
inline MultiOp swap(uint64_t m) { return detail::make(qvnt_op_swap, m); }
inline MultiOp sqrt_swap(uint64_t m) { return detail::make(qvnt_op_sqrt_swap, m); }
inline MultiOp i_swap(uint64_t m) { return detail::make(qvnt_op_i_swap, m); }
inline MultiOp sqrt_i_swap(uint64_t m) { return detail::make(qvnt_op_sqrt_i_swap, m); }
inline MultiOp h(uint64_t m) { return detail::make(qvnt_op_h, m); }
inline MultiOp u1(double lam, uint64_t m) { return detail::make(qvnt_op_u1, lam, m); }
inline MultiOp u2(double phi_, double lam, uint64_t m) { return detail::make(qvnt_op_u2, phi_, lam, m); }
inline MultiOp u3(double the, double phi_, double lam, uint64_t m) { return detail::make(qvnt_op_u3, the, phi_, lam, m); }
inline MultiOp qft(uint64_t m) { return detail::make(qvnt_op_qft, m); }
inline MultiOp qft_swapped(uint64_t m) { return detail::make(qvnt_op_qft_swapped, m); }
}  // namespace op

// CReg: a classical register of q_num bits (src/register/class.rs:50-155);
// measure_mask returns one with the measured bits in place (quant.rs:500).
class CReg {
public:
    CReg() = default;
    // class.rs:60-62, 66-74
    static CReg new_(uint32_t q_num) { return with_state(q_num, 0); }
    static CReg with_state(uint32_t q_num, uint64_t state) {
        CReg c;
        c.value_ = state;
        c.set_num(q_num);
        return c;
    }
    uint32_t num() const { return q_num_; }                         // class.rs:76-78
    void set_num(uint32_t q_num) {                                  // class.rs:80-83
        q_num_ = q_num;
        q_mask_ = q_num >= 64 ? ~0ull : (1ull << q_num) - 1u;       // wrapping_shl(q).wrapping_sub(1)
    }
    void reset(uint64_t i_state) { value_ = i_state & q_mask_; }    // class.rs:85-87
    void set(bool bit, uint64_t mask) { value_ = bit ? (value_ | mask) : (value_ & ~mask); }  // class.rs:89-95
    void xor_(bool bit, uint64_t mask) {                            // class.rs:97-101 (`xor` is a C++ keyword)
        if (bit) value_ ^= mask;
    }
    uint64_t get() const { return value_; }                         // class.rs:111-113
    // the bits of value at the set bits of mask & q_mask, packed from bit 0 (class.rs:115-125)
    uint64_t get_by_mask(uint64_t mask) const {
        uint64_t out = 0;
        unsigned k = 0;
        for (uint64_t m = mask & q_mask_; m; m &= m - 1, ++k)
            if (value_ & (m & (0 - m))) out |= 1ull << k;
        return out;
    }
    // tensor product, self on the low bits (class.rs:103-109, 142-154)
    CReg operator*(const CReg& o) const { return with_state(q_num_ + o.q_num_, value_ | (o.value_ << q_num_)); }
    CReg& operator*=(const CReg& o) { return *this = *this * o; }
    bool operator==(const CReg& o) const { return value_ == o.value_ && q_num_ == o.q_num_ && q_mask_ == o.q_mask_; }
    // Debug: "(b_{n-1} .. b_0)" (class.rs:128-140)
    std::string debug() const {
        std::string s;
        for (uint64_t m = q_mask_; m; m &= m - 1) s.insert(s.begin(), (value_ & (m & (0 - m))) ? '1' : '0');
        return "(" + s + ")";
    }

private:
    uint64_t value_ = 0;
    uint32_t q_num_ = 0;
    uint64_t q_mask_ = 0;
};

// VReg: single-bit masks of a register's qubits, indexed to build masks
// (src/register/virtl.rs:85-157).
class VReg {
public:
    static VReg new_(uint32_t num) { return from(num >= 64 ? ~0ull : (1ull << num) - 1u); }  // virtl.rs:88-91
    static VReg from(uint64_t mask) {                               // virtl.rs:93-104
        VReg v;
        for (uint64_t m = mask; m; m &= m - 1) v.bits_.push_back(m & (0 - m));
        return v;
    }
    size_t len() const { return bits_.size(); }
    uint64_t operator[](size_t idx) const { return bits_.at(idx); }  // virtl.rs:112-119
    // the OR of the masks at the positions where f holds (virtl.rs:121-137)
    template <class F, class = decltype(std::declval<F>()(size_t(0)))>
    uint64_t operator[](F f) const {
        uint64_t out = 0;
        for (size_t i = 0; i < bits_.size(); ++i)
            if (f(i)) out |= bits_[i];
        return out;
    }
    // positions listed (virtl.rs:139-145)
    uint64_t operator[](std::initializer_list<size_t> pos) const {
        return (*this)[([&](size_t i) { return std::find(pos.begin(), pos.end(), i) != pos.end(); })];
    }
    uint64_t all() const {                                          // v[..] (virtl.rs:147-157)
        return (*this)[([](size_t) { return true; })];
    }
    // Debug: the masks as zero-padded 32-digit hex (virtl.rs:106-110)
    std::string debug() const {
        std::string s = "[";
        char buf[40];
        for (size_t i = 0; i < bits_.size(); ++i) {
            std::snprintf(buf, sizeof buf, "%032llx", (unsigned long long)bits_[i]);
            s += (i ? ", " : "") + std::string(buf);
        }
        return s + "]";
    }

private:
    std::vector<uint64_t> bits_;
};

class QReg {
public:
    using C = std::complex<double>;
    static QReg new_(uint32_t q_num, qvnt_prec prec = QVNT_F64) { return QReg(q_num, 0, prec); }
    static QReg with_state(uint32_t q_num, uint64_t state, qvnt_prec prec = QVNT_F64) { return QReg(q_num, state, prec); }
    QReg(QReg&& o) noexcept : h_(std::exchange(o.h_, nullptr)), q_num_(o.q_num_), seed_(o.seed_), draws_(o.draws_) {}
    QReg(const QReg&) = delete;
    QReg& operator=(const QReg&) = delete;
    ~QReg() { qvnt_reg_free(h_); }

    uint32_t num() const { return q_num_; }
    uint64_t len() const { return qvnt_reg_len(h_); }
    // QReg::apply: one kernel per SingleOp
    QReg& apply(const MultiOp& o) {
        check(qvnt_reg_apply(h_, o.handle()));
        return *this;
    }
    // the fused form; returns the number of passes over the state
    uint32_t apply_fused(const MultiOp& o, uint32_t max_qubits = 0) {
        uint32_t passes = 0;
        check(qvnt_reg_apply_fused(h_, o.handle(), max_qubits, &passes));
        return passes;
    }
    double get_absolute() {
        double v = 0;
        check(qvnt_reg_get_absolute(h_, &v));
        return v;
    }
    std::vector<double> get_probabilities() {
        std::vector<double> p(size_t(1) << q_num_);
        check(qvnt_reg_get_probabilities(h_, p.data(), p.size()));
        return p;
    }
    // measure_mask with the register's next seeded draw, or a given one
    CReg measure_mask(uint64_t mask, std::optional<double> u = std::nullopt, bool renorm = false) {
        uint64_t out = 0;
        const double draw = u ? *u : qvnt_uniform(seed_, 0, draws_++);
        check(qvnt_reg_measure_mask(h_, mask, draw, renorm ? QVNT_MEASURE_RENORM : 0u, &out));
        return CReg::with_state(q_num_, out);
    }
    CReg measure(std::optional<double> u = std::nullopt) { return measure_mask(~0ull, u); }
    std::vector<C> get_state() {
        std::vector<C> s(len());
        check(qvnt_reg_get_state(h_, s.data(), s.size() * sizeof(C)));
        return s;
    }
    void set_state(const std::vector<C>& s) { check(qvnt_reg_set_state(h_, s.data(), s.size() * sizeof(C))); }
    void seed(uint64_t s) {
        seed_ = s;
        draws_ = 0;
    }
    // fused form: phases of >= min_ops primitives as dense 16x16 blocks (0: never)
    void set_dense(uint32_t min_ops) { check(qvnt_reg_set_dense(h_, min_ops)); }

    // reset and register algebra (src/register/quant.rs)
    void reset(uint64_t i_state = 0) { check(qvnt_reg_reset(h_, i_state)); }
    void reset_by_mask(uint64_t mask) { check(qvnt_reg_reset_by_mask(h_, mask)); }
    // Mul for QReg: the tensor product, *this on the low qubits
    QReg operator*(QReg& other) {
        qvnt_reg* out = nullptr;
        check(qvnt_reg_tensor(h_, other.h_, &out));
        return QReg(out, q_num_ + other.q_num_);
    }
    QReg combine(QReg& other) {
        qvnt_reg* out = nullptr;
        check(qvnt_reg_combine(h_, other.h_, &out));
        return QReg(out, q_num_ + 1);
    }
    QReg combine_with_unitary(QReg& other, const std::array<C, 4>& m) {
        qvnt_reg* out = nullptr;
        check(qvnt_reg_combine_with_unitary(h_, other.h_, reinterpret_cast<const double*>(m.data()), &out));
        return QReg(out, q_num_ + 1);
    }
    void linear_composition(QReg& other, C c0, C c1) {
        const double c[4] = {c0.real(), c0.imag(), c1.real(), c1.imag()};
        check(qvnt_reg_linear_composition(h_, other.h_, c));
    }
    void set_num(uint32_t q_num) {
        check(qvnt_reg_set_num(h_, q_num));
        q_num_ = q_num;
    }
    // (|psi_i|, arg psi_i) for i < 2^n
    std::vector<std::pair<double, double>> get_polar() {
        std::vector<std::pair<double, double>> out(size_t(1) << q_num_);
        check(qvnt_reg_get_polar(h_, reinterpret_cast<double*>(out.data()), 2 * out.size()));
        return out;
    }
    // exact multinomial histogram of `count` seeded shots, one count per buffer element
    std::vector<uint64_t> sample_all(uint64_t count, uint64_t stream = 1) {
        std::vector<uint64_t> hist(len());
        check(qvnt_reg_sample_all(h_, count, seed_, stream, hist.data(), hist.size()));
        return hist;
    }
    // QReg::get_vreg / get_vreg_by (quant.rs:232-243): nullopt for bits beyond the register
    VReg get_vreg() const { return VReg::new_(q_num_); }
    std::optional<VReg> get_vreg_by(uint64_t mask) const {
        const uint64_t q_mask = q_num_ >= 64 ? ~0ull : (1ull << q_num_) - 1u;
        if (mask & ~q_mask) return std::nullopt;
        return VReg::from(mask);
    }
    // Debug of QReg (quant.rs:603-623): "QReg { 0: Complex { re: .., im: .. }, .. }"
    std::string debug() {
        size_t need = 0;
        check(qvnt_reg_debug(h_, nullptr, 0, &need));
        std::string s(need, '\0');
        check(qvnt_reg_debug(h_, s.data(), s.size(), &need));
        s.resize(need - 1);
        return s;
    }
    qvnt_reg* handle() { return h_; }

private:
    QReg(uint32_t q_num, uint64_t state, qvnt_prec prec) : q_num_(q_num) { check(qvnt_reg_with_state(q_num, state, prec, &h_)); }
    QReg(qvnt_reg* h, uint32_t q_num) : h_(h), q_num_(q_num) {}
    qvnt_reg* h_ = nullptr;
    uint32_t q_num_ = 0;
    uint64_t seed_ = 0x51564E54;
    uint64_t draws_ = 0;
};

}  // namespace qvnt
