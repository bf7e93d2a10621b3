// qvnt.hpp -- C++17 facade over the C ABI (qvnt.h) with the shape of
// qvnt 0.4.4's Rust API:
//   qvnt::op::{id,x,y,z,s,t,rx,ry,rz,rxx,ryy,rzz,phi,swap,sqrt_swap,i_swap,
//              sqrt_i_swap,h,u1,u2,u3,qft,qft_swapped}   src/operator/mod.rs:113-522
//   MultiOp: `a * b` appends, `.c(mask)` -> std::optional (the reference's
//            Option), `.dgr()`, `.act_on()`, Debug name    src/operator/multi/mod.rs
//   QReg::{new_, with_state, apply, apply_fused, get_probabilities,
//          get_absolute, measure_mask, measure}            src/register/quant.rs
//   CReg (measurement result)                               src/register/class.rs
// Where the reference panics (a rotation mask with the wrong bit count, a
// measurement of a zero state) the facade throws qvnt::Error.
#pragma once

#include <array>
#include <complex>
#include <cstdint>
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

// CReg: the measured value, bits in place (class.rs:66-74, quant.rs:500).
struct CReg {
    uint32_t q_num = 0;
    uint64_t value = 0;
    uint64_t get() const { return value; }
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
        return CReg{q_num_, out};
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
