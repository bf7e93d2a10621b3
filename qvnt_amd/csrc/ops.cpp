// Host-side operator algebra: the op:: constructors, MultiOp queue
// operations (`*`, .c(), .dgr()) and the Debug names, behind the C ABI.
//
// Semantics follow the reference (paths relative to its root):
//   constructors      src/operator/mod.rs:113-522, single/{pauli,rotate,swap}.rs
//   stored parameters src/operator/atomic/<gate>.rs `fn new`
//   h / qft           src/operator/multi/h.rs:14-45, multi/qft.rs:4-52
//   c / dgr / `*`     src/operator/multi/mod.rs:96-191, single/mod.rs:83-119
//   names             single/mod.rs:70-76 and each atomic `fn name`
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstring>
#include <new>

#include "common.hpp"

namespace qvnt {

static thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
void clear_error() { g_last_error.clear(); }

bool is_diagonal_kind(uint32_t kind) {
    switch (kind) {
    case QVNT_ID: case QVNT_Z: case QVNT_S: case QVNT_T:
    case QVNT_RZ: case QVNT_RZZ: case QVNT_PHI:
        return true;
    default:
        return false;
    }
}

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kHalfPi = 1.57079632679489661923;

qvnt_gate make(uint32_t kind, uint64_t act) {
    qvnt_gate g;
    std::memset(&g, 0, sizeof g);
    g.kind = kind;
    g.act = act;
    g.a = act;
    return g;
}

// (cos x, sin x) from glibc's sincos.  The reference evaluates
// `C::new(x.cos(), x.sin())` (atomic/rx.rs:10-14 and siblings); an optimised
// Rust build on linux-gnu lowers that pair of the same argument to one
// sincos call, whose last bit differs from separate cos/sin calls for about
// 0.1% of arguments.
void cos_sin(double x, double* c, double* s) { ::sincos(x, s, c); }

// A rotation stores (cos th/2, sin th/2).
int rotation(uint32_t kind, double theta, uint64_t mask, unsigned bits, qvnt_gate* out) {
    if (popc(mask) != bits)
        return fail(QVNT_E_INVALID_MASK, bits == 1 ? "Mask should contain 1 bit!" : "Mask should contain 2 bit!");
    *out = make(kind, mask);
    const double half = theta * 0.5;  // identical to theta / 2 for every finite theta
    cos_sin(half, &out->re, &out->im);
    return QVNT_OK;
}

int emit(std::vector<qvnt_gate> gates, qvnt_op** out) {
    if (!out) return fail(QVNT_E_ARG, "null output pointer");
    qvnt_op* op = new (std::nothrow) qvnt_op;
    if (!op) return fail(QVNT_E_OOM, "host allocation failed");
    op->g = std::move(gates);
    *out = op;
    clear_error();
    return QVNT_OK;
}

int single(uint32_t kind, uint64_t mask, qvnt_op** out) { return emit({make(kind, mask)}, out); }

int single_rot(uint32_t kind, double theta, uint64_t mask, unsigned bits, qvnt_op** out) {
    qvnt_gate g;
    int rc = rotation(kind, theta, mask, bits, &g);
    if (rc) return rc;
    return emit({g}, out);
}

int single_2bit(uint32_t kind, uint64_t mask, qvnt_op** out) {
    if (popc(mask) != 2) return fail(QVNT_E_INVALID_MASK, "Mask should contain 2 bit!");
    return emit({make(kind, mask)}, out);
}

// H^{(x)k} on the set bits: consecutive set bits (ascending) pair into one
// 2-qubit H2 whose `a` is the higher bit and `b` the lower one; an odd bit
// left at the end becomes an H1 (multi/h.rs:14-45).
void hadamard(uint64_t mask, std::vector<qvnt_gate>& q) {
    const unsigned k = popc(mask);
    if (k == 0) return;
    if (k == 1) { q.push_back(make(QVNT_H1, mask)); return; }
    uint64_t pending = 0;
    for (uint64_t rest = mask; rest; rest &= rest - 1) {
        const uint64_t bit = rest & (~rest + 1);
        if (!pending) { pending = bit; continue; }
        qvnt_gate g = make(QVNT_H2, bit | pending);
        g.a = bit;
        g.b = pending;
        q.push_back(g);
        pending = 0;
    }
    if (pending) q.push_back(make(QVNT_H1, pending));
}

// The reference's QFT ladder (multi/qft.rs:4-33): bits ascending; for every
// bit but the last, H on it, then RZ(pi / 2^j) on the j-th next bit
// controlled by it; H on the last bit.
void qft_ladder(uint64_t mask, std::vector<qvnt_gate>& q) {
    const unsigned k = popc(mask);
    if (k <= 1) { hadamard(mask, q); return; }
    uint64_t bits[64];
    unsigned m = 0;
    for (uint64_t rest = mask; rest; rest &= rest - 1) bits[m++] = rest & (~rest + 1);
    for (unsigned i = 0; i + 1 < k; ++i) {
        q.push_back(make(QVNT_H1, bits[i]));
        for (unsigned j = 1; i + j < k; ++j) {
            qvnt_gate g;
            rotation(QVNT_RZ, std::ldexp(kPi, -(int)j), bits[i + j], 1, &g);
            g.ctrl = bits[i];
            q.push_back(g);
        }
    }
    q.push_back(make(QVNT_H1, bits[k - 1]));
}

// ------------------------------------------------------------ unitarity
// The reference accepts a matrix when float-cmp approx_eq (2 ulps or
// f64::EPSILON absolute) holds for the row norms and for re+im of the row
// inner products (math/matrix.rs:23-66, approx_cmp.rs:5-8).
bool approx_eq(double x, double y) {
    if (std::fabs(x - y) <= 2.220446049250313e-16) return true;
    if ((x < 0) != (y < 0)) return x == y;
    int64_t ix, iy;
    std::memcpy(&ix, &x, 8);
    std::memcpy(&iy, &y, 8);
    return (ix > iy ? ix - iy : iy - ix) <= 2;
}
struct cd { double re, im; };
cd at(const double* m, int k) { return {m[2 * k], m[2 * k + 1]}; }
cd mul_conj(cd a, cd b) { return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im}; }  // a * conj(b)
cd add(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }
double nsq(cd a) { return a.re * a.re + a.im * a.im; }

bool unitary_1(const double* m) {
    const double r0 = nsq(at(m, 0)) + nsq(at(m, 1));
    const double r1 = nsq(at(m, 2)) + nsq(at(m, 3));
    const cd x = add(mul_conj(at(m, 0), at(m, 2)), mul_conj(at(m, 1), at(m, 3)));
    return approx_eq(r0, 1.0) && approx_eq(r1, 1.0) && approx_eq(x.re + x.im, 0.0);
}
bool unitary_2(const double* m) {
    for (int i = 0; i < 4; ++i) {
        const double r = (nsq(at(m, 4 * i)) + nsq(at(m, 4 * i + 1))) + (nsq(at(m, 4 * i + 2)) + nsq(at(m, 4 * i + 3)));
        if (!approx_eq(r, 1.0)) return false;
    }
    for (int i = 0; i < 4; ++i)
        for (int j = i + 1; j < 4; ++j) {
            const cd x = add(add(mul_conj(at(m, 4 * i), at(m, 4 * j)), mul_conj(at(m, 4 * i + 1), at(m, 4 * j + 1))),
                             add(mul_conj(at(m, 4 * i + 2), at(m, 4 * j + 2)), mul_conj(at(m, 4 * i + 3), at(m, 4 * j + 3))));
            if (!approx_eq(x.re + x.im, 0.0)) return false;
        }
    return true;
}

void conj_transpose(double* m, int dim) {
    double t[32];
    for (int r = 0; r < dim; ++r)
        for (int c = 0; c < dim; ++c) {
            t[2 * (dim * r + c)] = m[2 * (dim * c + r)];
            t[2 * (dim * r + c) + 1] = -m[2 * (dim * c + r) + 1];
        }
    std::memcpy(m, t, sizeof(double) * 2 * dim * dim);
}

// The inverse each atomic op reports (atomic/<gate>.rs `fn dgr`).  Note the
// rotations: the reference negates the stored phase, which yields -U rather
// than U^dagger (rx.rs:44 etc.); that behaviour is kept.
void dagger(qvnt_gate& g) {
    switch (g.kind) {
    case QVNT_S: case QVNT_T: case QVNT_ISWAP: case QVNT_SQRT_SWAP: case QVNT_SQRT_ISWAP:
        g.dagger ^= 1u;
        break;
    case QVNT_RX: case QVNT_RY: case QVNT_RZ: case QVNT_RXX: case QVNT_RYY: case QVNT_RZZ:
        g.re = -g.re;
        g.im = -g.im;
        break;
    case QVNT_PHI:
        g.im = -g.im;
        break;
    case QVNT_U1: conj_transpose(g.m, 2); break;
    case QVNT_U2: conj_transpose(g.m, 4); break;
    default: break;
    }
}

// --------------------------------------------------------------- names
// Rust's `{}` (Display) and `{:?}` (Debug) of f64, from the shortest
// round-trip digits.
struct Digits {
    bool neg = false;
    std::string d;  // significant digits, no dot
    int e = 0;      // decimal exponent of the first digit
};
Digits shortest(double v) {
    char buf[64];
    auto res = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::scientific);
    std::string s(buf, res.ptr);
    Digits out;
    size_t p = 0;
    if (s[0] == '-') { out.neg = true; p = 1; }
    size_t epos = s.find('e');
    for (size_t i = p; i < epos; ++i)
        if (s[i] != '.') out.d.push_back(s[i]);
    out.e = std::stoi(s.substr(epos + 1));
    while (out.d.size() > 1 && out.d.back() == '0') out.d.pop_back();
    return out;
}
std::string plain(const Digits& x, bool force_fraction) {
    std::string ip, fp;
    if (x.e >= 0) {
        for (int i = 0; i <= x.e; ++i) ip.push_back(i < (int)x.d.size() ? x.d[i] : '0');
        if ((int)x.d.size() > x.e + 1) fp = x.d.substr(x.e + 1);
    } else {
        ip = "0";
        fp = std::string(-x.e - 1, '0') + x.d;
    }
    if (fp == "0" || (x.d == "0")) fp.clear();
    if (fp.empty() && force_fraction) fp = "0";
    return (x.neg ? "-" : "") + ip + (fp.empty() ? "" : "." + fp);
}
std::string rust_display(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v < 0 ? "-inf" : "inf";
    return plain(shortest(v), false);
}
std::string rust_debug(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v < 0 ? "-inf" : "inf";
    const double a = std::fabs(v);
    Digits x = shortest(v);
    if ((a != 0.0 && a < 1e-4) || a >= 1e16) {
        std::string m = x.d.substr(0, 1);
        if (x.d.size() > 1) m += "." + x.d.substr(1);
        return (x.neg ? "-" : "") + m + "e" + std::to_string(x.e);
    }
    return plain(x, true);
}
std::string complex_debug(const double* m, int k) {
    return "Complex { re: " + rust_debug(m[2 * k]) + ", im: " + rust_debug(m[2 * k + 1]) + " }";
}
std::string angle(const qvnt_gate& g) { return rust_display(2.0 * std::atan2(g.im, g.re)); }

std::string gate_name(const qvnt_gate& g) {
    const std::string a = std::to_string(g.a);
    std::string s;
    switch (g.kind) {
    case QVNT_ID: s = "Id"; break;
    case QVNT_X: s = "X" + a; break;
    case QVNT_Y: s = "Y" + a; break;
    case QVNT_Z: s = "Z" + a; break;
    case QVNT_S: s = "S" + a; break;
    case QVNT_T: s = "T" + a; break;
    case QVNT_RX: s = "RX" + a + "(" + angle(g) + ")"; break;
    case QVNT_RY: s = "RY" + a + "(" + angle(g) + ")"; break;
    case QVNT_RZ: s = "RZ" + a + "(" + angle(g) + ")"; break;
    case QVNT_RXX: s = "RXX" + a + "(" + angle(g) + ")"; break;
    case QVNT_RYY: s = "RYY" + a + "(" + angle(g) + ")"; break;
    case QVNT_RZZ: s = "RZZ" + a + "(" + angle(g) + ")"; break;
    case QVNT_PHI: s = "PHI" + a + "(" + rust_display(std::atan2(g.im, g.re)) + ")"; break;
    case QVNT_H1: s = "H" + a; break;
    case QVNT_H2: s = "H" + std::to_string(g.a | g.b); break;
    case QVNT_SWAP: s = "SWAP" + a; break;
    case QVNT_ISWAP: s = "iSWAP" + a; break;
    case QVNT_SQRT_SWAP: s = "sqrt(SWAP" + a + ")"; break;
    case QVNT_SQRT_ISWAP: s = "sqrt(iSWAP" + a + ")"; break;
    case QVNT_U1:
        s = "U" + a + "[[" + complex_debug(g.m, 0) + ", " + complex_debug(g.m, 1) + "], [" +
            complex_debug(g.m, 2) + ", " + complex_debug(g.m, 3) + "]]";
        break;
    case QVNT_U2: {
        s = "U" + std::to_string(g.a | g.b) + "[";
        for (int r = 0; r < 4; ++r) {
            s += r ? ", [" : "[";
            for (int c = 0; c < 4; ++c) s += (c ? ", " : "") + complex_debug(g.m, 4 * r + c);
            s += "]";
        }
        s += "]";
        break;
    }
    default: s = "?"; break;
    }
    if (g.ctrl) s = "C" + std::to_string(g.ctrl) + "_" + s;
    return s;
}

}  // namespace

// num-complex's Debug of a Complex<f64>: "Complex { re: .., im: .. }".
std::string complex_debug_value(double re, double im) {
    const double m[2] = {re, im};
    return complex_debug(m, 0);
}

bool is_unitary_m1(const double* m) { return unitary_1(m); }

}  // namespace qvnt

using namespace qvnt;

extern "C" {

const char* qvnt_last_error(void) { return g_last_error.c_str(); }
const char* qvnt_version(void) { return "qvnt-mi355x 0.1.0 (gfx950)"; }

double qvnt_uniform(uint64_t seed, uint64_t stream, uint64_t index) {
    auto mix = [](uint64_t z) {
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    const uint64_t key = mix(seed ^ mix(stream ^ 0x51564E54ull));
    return (double)(mix(key + (index + 1u) * 0x9E3779B97F4A7C15ull) >> 11) * 0x1.0p-53;
}

int qvnt_op_id(qvnt_op** out) { return emit({}, out); }
int qvnt_op_x(uint64_t a, qvnt_op** out) { return single(QVNT_X, a, out); }
int qvnt_op_y(uint64_t a, qvnt_op** out) { return single(QVNT_Y, a, out); }
int qvnt_op_z(uint64_t a, qvnt_op** out) { return single(QVNT_Z, a, out); }
int qvnt_op_s(uint64_t a, qvnt_op** out) { return single(QVNT_S, a, out); }
int qvnt_op_t(uint64_t a, qvnt_op** out) { return single(QVNT_T, a, out); }
int qvnt_op_rx(double p, uint64_t a, qvnt_op** out) { return single_rot(QVNT_RX, p, a, 1, out); }
int qvnt_op_ry(double p, uint64_t a, qvnt_op** out) { return single_rot(QVNT_RY, p, a, 1, out); }
int qvnt_op_rz(double p, uint64_t a, qvnt_op** out) { return single_rot(QVNT_RZ, p, a, 1, out); }
int qvnt_op_rxx(double p, uint64_t ab, qvnt_op** out) { return single_rot(QVNT_RXX, p, ab, 2, out); }
int qvnt_op_ryy(double p, uint64_t ab, qvnt_op** out) { return single_rot(QVNT_RYY, p, ab, 2, out); }
int qvnt_op_rzz(double p, uint64_t ab, qvnt_op** out) { return single_rot(QVNT_RZZ, p, ab, 2, out); }
int qvnt_op_phi(double p, uint64_t a, qvnt_op** out) {
    if (popc(a) != 1) return fail(QVNT_E_INVALID_MASK, "Mask should contain 1 bit!");
    qvnt_gate g = make(QVNT_PHI, a);
    cos_sin(p, &g.re, &g.im);
    return emit({g}, out);
}
int qvnt_op_swap(uint64_t ab, qvnt_op** out) { return single_2bit(QVNT_SWAP, ab, out); }
int qvnt_op_sqrt_swap(uint64_t ab, qvnt_op** out) { return single_2bit(QVNT_SQRT_SWAP, ab, out); }
int qvnt_op_i_swap(uint64_t ab, qvnt_op** out) { return single_2bit(QVNT_ISWAP, ab, out); }
int qvnt_op_sqrt_i_swap(uint64_t ab, qvnt_op** out) { return single_2bit(QVNT_SQRT_ISWAP, ab, out); }
int qvnt_op_h(uint64_t a, qvnt_op** out) {
    std::vector<qvnt_gate> q;
    hadamard(a, q);
    return emit(std::move(q), out);
}
int qvnt_op_u1(double lam, uint64_t a, qvnt_op** out) { return qvnt_op_rz(lam, a, out); }
static int u_seq(double first_rz, double ry_angle, double last_rz, uint64_t a, qvnt_op** out) {
    qvnt_gate g[3];
    int rc;
    if ((rc = rotation(QVNT_RZ, first_rz, a, 1, &g[0]))) return rc;
    if ((rc = rotation(QVNT_RY, ry_angle, a, 1, &g[1]))) return rc;
    if ((rc = rotation(QVNT_RZ, last_rz, a, 1, &g[2]))) return rc;
    return emit({g[0], g[1], g[2]}, out);
}
int qvnt_op_u2(double phi, double lam, uint64_t a, qvnt_op** out) { return u_seq(lam, kHalfPi, phi, a, out); }
int qvnt_op_u3(double the, double phi, double lam, uint64_t a, qvnt_op** out) { return u_seq(lam, the, phi, a, out); }
int qvnt_op_qft(uint64_t a, qvnt_op** out) {
    std::vector<qvnt_gate> q;
    qft_ladder(a, q);
    return emit(std::move(q), out);
}
int qvnt_op_qft_swapped(uint64_t a, qvnt_op** out) {
    std::vector<qvnt_gate> q;
    qft_ladder(a, q);
    uint64_t bits[64];
    unsigned m = 0;
    for (uint64_t rest = a; rest; rest &= rest - 1) bits[m++] = rest & (~rest + 1);
    for (unsigned i = 0; i < m / 2; ++i) q.push_back(make(QVNT_SWAP, bits[i] | bits[m - 1 - i]));
    return emit(std::move(q), out);
}
int qvnt_op_matrix1(uint64_t a, const double m[8], qvnt_op** out) {
    if (!m) return fail(QVNT_E_ARG, "null matrix");
    if (popc(a) != 1 || !unitary_1(m)) return fail(QVNT_E_INVALID_MASK, "invalid 1-qubit unitary or mask");
    qvnt_gate g = make(QVNT_U1, a);
    std::memcpy(g.m, m, 8 * sizeof(double));
    return emit({g}, out);
}
int qvnt_op_matrix2(uint64_t a, uint64_t b, const double m[32], qvnt_op** out) {
    if (!m) return fail(QVNT_E_ARG, "null matrix");
    if (popc(a) != 1 || popc(b) != 1 || !unitary_2(m))
        return fail(QVNT_E_INVALID_MASK, "invalid 2-qubit unitary or mask");
    qvnt_gate g = make(QVNT_U2, a);  // acts_on() reports a only (atomic/u2.rs:70-72)
    g.b = b;
    std::memcpy(g.m, m, 32 * sizeof(double));
    return emit({g}, out);
}
int qvnt_op_from_gates(const qvnt_gate* gates, size_t count, qvnt_op** out) {
    if (count && !gates) return fail(QVNT_E_ARG, "null gate array");
    std::vector<qvnt_gate> q;
    q.reserve(count);
    for (size_t i = 0; i < count; ++i) {
        if (gates[i].kind >= QVNT_KIND_COUNT) return fail(QVNT_E_ARG, "unknown gate kind");
        if (gates[i].kind == QVNT_ID && gates[i].ctrl == 0) continue;  // name "Id" dropped
        q.push_back(gates[i]);
    }
    return emit(std::move(q), out);
}
int qvnt_op_clone(const qvnt_op* op, qvnt_op** out) {
    if (!op) return fail(QVNT_E_ARG, "null op");
    return emit(op->g, out);
}
void qvnt_op_free(qvnt_op* op) { delete op; }
int qvnt_op_mul(qvnt_op* lhs, const qvnt_op* rhs) {
    if (!lhs || !rhs) return fail(QVNT_E_ARG, "null op");
    if (lhs == rhs) {
        std::vector<qvnt_gate> copy = rhs->g;
        lhs->g.insert(lhs->g.end(), copy.begin(), copy.end());
    } else {
        lhs->g.insert(lhs->g.end(), rhs->g.begin(), rhs->g.end());
    }
    return QVNT_OK;
}
uint64_t qvnt_op_act_on(const qvnt_op* op) {
    uint64_t m = 0;
    if (op)
        for (const qvnt_gate& g : op->g) m |= g.act | g.ctrl;
    return m;
}
int qvnt_op_c(qvnt_op* op, uint64_t c) {
    if (!op) return fail(QVNT_E_ARG, "null op");
    if (qvnt_op_act_on(op) & c) return fail(QVNT_E_CTRL_OVERLAP, "control mask overlaps the operator");
    for (qvnt_gate& g : op->g) g.ctrl |= c;
    return QVNT_OK;
}
int qvnt_op_dgr(qvnt_op* op) {
    if (!op) return fail(QVNT_E_ARG, "null op");
    for (qvnt_gate& g : op->g) dagger(g);
    std::reverse(op->g.begin(), op->g.end());
    return QVNT_OK;
}
size_t qvnt_op_len(const qvnt_op* op) { return op ? op->g.size() : 0; }
const qvnt_gate* qvnt_op_gates(const qvnt_op* op) { return op && !op->g.empty() ? op->g.data() : nullptr; }
int qvnt_op_name(const qvnt_op* op, char* buf, size_t cap, size_t* need) {
    if (!op) return fail(QVNT_E_ARG, "null op");
    std::string s = "[";
    for (size_t i = 0; i < op->g.size(); ++i) s += (i ? ", " : "") + gate_name(op->g[i]);
    s += "]";
    if (need) *need = s.size() + 1;
    if (buf && cap) {
        const size_t n = std::min(cap - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = '\0';
    }
    return QVNT_OK;
}

}  // extern "C"
