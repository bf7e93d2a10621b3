// Register algebra, reset and readback (src/register/quant.rs:156-166,
// 202-229, 245-371, 397-431, 513-594) and the QASM driver loop
// (src/qasm/sym.rs:40-74) on device registers.
//
// Every kernel is a grid-stride stream over the output buffer (HBM-bound:
// one read of each input element and one write of each output element).
// FP contraction is off, so complex products and sums round exactly as the
// reference's num-complex arithmetic does.
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <vector>

#include "device.hpp"
#include "reg.hpp"

namespace qvnt {
namespace {

constexpr int kBlock = 256;
constexpr uint64_t kMaxGrid = 1u << 20;

dim3 grid_for(uint64_t n) {
    const uint64_t nb = (n + kBlock - 1) / kBlock;
    return dim3((unsigned)(nb < kMaxGrid ? nb : kMaxGrid));
}

#define QVNT_GRID_STRIDE(i, n) \
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * kBlock)

// num-complex Mul / Add (a * b = (a.re b.re - a.im b.im, a.re b.im + a.im b.re))
template <class R>
__device__ __forceinline__ cplx<R> cm(cplx<R> a, cplx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class R>
__device__ __forceinline__ cplx<R> ca(cplx<R> a, cplx<R> b) {
    return {a.re + b.re, a.im + b.im};
}

// tensor_prod (quant.rs:330-371): out[i] = a[i & ma] * b[(i >> na) & mb]
// for i < 2^(na+nb), zero in the padding.
template <class R>
__global__ __launch_bounds__(kBlock) void k_tensor(cplx<R>* __restrict__ out, uint64_t len, uint64_t q_size,
                                                   const cplx<R>* __restrict__ a, uint64_t ma, uint32_t na,
                                                   const cplx<R>* __restrict__ b, uint64_t mb) {
    QVNT_GRID_STRIDE(i, len) out[i] = i < q_size ? cm(a[i & ma], b[(i >> na) & mb]) : cplx<R>{(R)0, (R)0};
}

// combine (quant.rs:245-268): a then b, back to back.
template <class R>
__global__ __launch_bounds__(kBlock) void k_combine(cplx<R>* __restrict__ out, const cplx<R>* __restrict__ a,
                                                    const cplx<R>* __restrict__ b, uint64_t half) {
    QVNT_GRID_STRIDE(i, 2 * half) out[i] = i < half ? a[i] : b[i - half];
}

// combine_with_unitary (quant.rs:271-308).
template <class R>
__global__ __launch_bounds__(kBlock) void k_combine_u(cplx<R>* __restrict__ out, uint64_t len, const cplx<R>* __restrict__ a,
                                                      const cplx<R>* __restrict__ b, uint64_t m, cplx<R> c00,
                                                      cplx<R> c01, cplx<R> c10, cplx<R> c11) {
    QVNT_GRID_STRIDE(i, len) {
        const cplx<R> q0 = a[m & i], q1 = b[m & i];
        out[i] = (~m & i) == 0 ? ca(cm(c00, q0), cm(c01, q1)) : ca(cm(c10, q0), cm(c11, q1));
    }
}

// linear_composition (quant.rs:311-328): psi = psi * c0 + other * c1.
template <class R>
__global__ __launch_bounds__(kBlock) void k_lincomb(cplx<R>* __restrict__ psi, const cplx<R>* __restrict__ other,
                                                    uint64_t len, cplx<R> c0, cplx<R> c1) {
    QVNT_GRID_STRIDE(i, len) psi[i] = ca(cm(psi[i], c0), cm(other[i], c1));
}

// set_num (quant.rs:156-166): keep the first `keep` amplitudes, zero the rest.
template <class R>
__global__ __launch_bounds__(kBlock) void k_resize(cplx<R>* __restrict__ out, uint64_t len, const cplx<R>* __restrict__ in,
                                                   uint64_t keep) {
    QVNT_GRID_STRIDE(i, len) out[i] = i < keep ? in[i] : cplx<R>{(R)0, (R)0};
}

// get_polar (quant.rs:417-431): num-complex to_polar = (hypot(re, im), atan2(im, re)), in FP64.
template <class R>
__global__ __launch_bounds__(kBlock) void k_polar(const cplx<R>* __restrict__ psi, uint64_t n, double* __restrict__ out) {
    QVNT_GRID_STRIDE(i, n) {
        const double re = (double)psi[i].re, im = (double)psi[i].im;
        out[2 * i] = hypot(re, im);
        out[2 * i + 1] = atan2(im, re);
    }
}

// `waiter` runs after everything queued so far on `producer`.
int order_after(hipStream_t waiter, hipStream_t producer) {
    if (waiter == producer) return QVNT_OK;
    hipEvent_t e;
    QVNT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    hipError_t err = hipEventRecord(e, producer);
    if (err == hipSuccess) err = hipStreamWaitEvent(waiter, e, 0);
    (void)hipEventDestroy(e);
    if (err != hipSuccess) return fail(QVNT_E_HIP, std::string("stream ordering: ") + hipGetErrorString(err));
    return QVNT_OK;
}

int check_pair(const qvnt_reg* a, const qvnt_reg* b) {
    if (!a || !b) return fail(QVNT_E_ARG, "null register");
    if (a->prec != b->prec) return fail(QVNT_E_ARG, "registers of different precision");
    return QVNT_OK;
}

// A fresh register of q_num qubits in a's precision; the caller fills it.
int fresh(const qvnt_reg* like, uint32_t q_num, qvnt_reg** out) {
    return qvnt_reg_with_state(q_num, 0, like->prec, out);
}

template <class R>
cplx<R> coef(const double* c, int k) {
    return cplx<R>{(R)c[2 * k], (R)c[2 * k + 1]};
}

// Launches `body(out)` on out's stream after a and b, and keeps a and b
// from being overwritten before it has read them.
template <class F>
int run_binary(qvnt_reg* a, qvnt_reg* b, qvnt_reg* out, F&& body) {
    if (int rc = order_after(out->stream, a->stream)) return rc;
    if (int rc = order_after(out->stream, b->stream)) return rc;
    if (int rc = body()) return rc;
    if (int rc = order_after(a->stream, out->stream)) return rc;
    return order_after(b->stream, out->stream);
}

// QReg::normalize (quant.rs:397-414) over the whole buffer.
int normalize(qvnt_reg* r) {
    double total = 0.0;
    if (int rc = tree_build(r, r->len, &total)) return rc;
    double norm = std::sqrt(total);
    if (norm <= 1e-15) return launch_init(r, 0);
    if (1. - norm <= 1e-9) return QVNT_OK;
    return rescale(r, 1. / norm);
}

uint64_t get_by_mask(uint64_t value, uint64_t mask) {  // class.rs:106-117
    uint64_t out = 0;
    unsigned k = 0;
    for (uint64_t m = mask; m; m &= m - 1, ++k)
        if (value & (m & (~m + 1))) out |= 1ull << k;
    return out;
}

}  // namespace
}  // namespace qvnt

using namespace qvnt;

extern "C" {

int qvnt_reg_reset(qvnt_reg* r, uint64_t i_state) {
    if (!r) return fail(QVNT_E_ARG, "null register");
    return launch_init(r, i_state & ((1ull << r->q_num) - 1u));
}

int qvnt_reg_reset_by_mask(qvnt_reg* r, uint64_t mask) {
    if (!r) return fail(QVNT_E_ARG, "null register");
    const uint64_t q_mask = (1ull << r->q_num) - 1u;
    if ((mask & q_mask) == q_mask) return launch_init(r, 0);
    if (int rc = collapse(r, 0, mask, false)) return rc;  // zero idx & mask != 0
    return normalize(r);
}

int qvnt_reg_tensor(qvnt_reg* a, qvnt_reg* b, qvnt_reg** out) {
    if (int rc = check_pair(a, b)) return rc;
    if (!out) return fail(QVNT_E_ARG, "null output pointer");
    const uint32_t q_num = a->q_num + b->q_num;
    if (q_num > 40) return fail(QVNT_E_OOM, "register too large");
    qvnt_reg* r = nullptr;
    if (int rc = fresh(a, q_num, &r)) return rc;
    const uint64_t q_size = 1ull << q_num, ma = (1ull << a->q_num) - 1u, mb = (1ull << b->q_num) - 1u;
    int rc = run_binary(a, b, r, [&]() -> int {
        if (r->prec == QVNT_F64)
            hipLaunchKernelGGL(k_tensor<double>, grid_for(r->len), dim3(kBlock), 0, r->stream, static_cast<c64*>(r->psi),
                               r->len, q_size, static_cast<const c64*>(a->psi), ma, a->q_num,
                               static_cast<const c64*>(b->psi), mb);
        else
            hipLaunchKernelGGL(k_tensor<float>, grid_for(r->len), dim3(kBlock), 0, r->stream, static_cast<c32*>(r->psi),
                               r->len, q_size, static_cast<const c32*>(a->psi), ma, a->q_num,
                               static_cast<const c32*>(b->psi), mb);
        QVNT_HIP(hipGetLastError());
        return QVNT_OK;
    });
    if (rc) {
        qvnt_reg_free(r);
        return rc;
    }
    *out = r;
    return QVNT_OK;
}

int qvnt_reg_combine(qvnt_reg* a, qvnt_reg* b, qvnt_reg** out) {
    if (int rc = check_pair(a, b)) return rc;
    if (!out) return fail(QVNT_E_ARG, "null output pointer");
    if (a->q_num != b->q_num) return fail(QVNT_E_ARG, "combine needs registers of equal size");
    if (a->q_num < 3) return fail(QVNT_E_ARG, "combine needs at least 3 qubits (below, the 8-element buffers do not fit as halves)");
    if (a->q_num + 1 > 40) return fail(QVNT_E_OOM, "register too large");
    qvnt_reg* r = nullptr;
    if (int rc = fresh(a, a->q_num + 1, &r)) return rc;
    int rc = run_binary(a, b, r, [&]() -> int {
        if (r->prec == QVNT_F64)
            hipLaunchKernelGGL(k_combine<double>, grid_for(r->len), dim3(kBlock), 0, r->stream, static_cast<c64*>(r->psi),
                               static_cast<const c64*>(a->psi), static_cast<const c64*>(b->psi), a->len);
        else
            hipLaunchKernelGGL(k_combine<float>, grid_for(r->len), dim3(kBlock), 0, r->stream, static_cast<c32*>(r->psi),
                               static_cast<const c32*>(a->psi), static_cast<const c32*>(b->psi), a->len);
        QVNT_HIP(hipGetLastError());
        return QVNT_OK;
    });
    if (rc) {
        qvnt_reg_free(r);
        return rc;
    }
    *out = r;
    return QVNT_OK;
}

int qvnt_reg_combine_with_unitary(qvnt_reg* a, qvnt_reg* b, const double m[8], qvnt_reg** out) {
    if (int rc = check_pair(a, b)) return rc;
    if (!out || !m) return fail(QVNT_E_ARG, "null argument");
    if (a->q_num != b->q_num) return fail(QVNT_E_ARG, "combine needs registers of equal size");
    if (a->q_num + 1 > 40) return fail(QVNT_E_OOM, "register too large");
    if (!is_unitary_m1(m)) return fail(QVNT_E_ARG, "combine_with_unitary needs a unitary 2x2 matrix");
    qvnt_reg* r = nullptr;
    if (int rc = fresh(a, a->q_num + 1, &r)) return rc;
    const uint64_t qm = (1ull << a->q_num) - 1u;
    int rc = run_binary(a, b, r, [&]() -> int {
        if (r->prec == QVNT_F64)
            hipLaunchKernelGGL(k_combine_u<double>, grid_for(r->len), dim3(kBlock), 0, r->stream, static_cast<c64*>(r->psi),
                               r->len, static_cast<const c64*>(a->psi), static_cast<const c64*>(b->psi), qm,
                               coef<double>(m, 0), coef<double>(m, 1), coef<double>(m, 2), coef<double>(m, 3));
        else
            hipLaunchKernelGGL(k_combine_u<float>, grid_for(r->len), dim3(kBlock), 0, r->stream, static_cast<c32*>(r->psi),
                               r->len, static_cast<const c32*>(a->psi), static_cast<const c32*>(b->psi), qm,
                               coef<float>(m, 0), coef<float>(m, 1), coef<float>(m, 2), coef<float>(m, 3));
        QVNT_HIP(hipGetLastError());
        return QVNT_OK;
    });
    if (rc) {
        qvnt_reg_free(r);
        return rc;
    }
    *out = r;
    return QVNT_OK;
}

int qvnt_reg_linear_composition(qvnt_reg* r, qvnt_reg* other, const double c[4]) {
    if (int rc = check_pair(r, other)) return rc;
    if (!c) return fail(QVNT_E_ARG, "null coefficients");
    if (r->len != other->len) return fail(QVNT_E_ARG, "linear_composition needs buffers of equal length");
    if (r == other) return fail(QVNT_E_ARG, "linear_composition of a register with itself");
    if (int rc = order_after(r->stream, other->stream)) return rc;
    if (r->prec == QVNT_F64)
        hipLaunchKernelGGL(k_lincomb<double>, grid_for(r->len), dim3(kBlock), 0, r->stream, static_cast<c64*>(r->psi),
                           static_cast<const c64*>(other->psi), r->len, coef<double>(c, 0), coef<double>(c, 1));
    else
        hipLaunchKernelGGL(k_lincomb<float>, grid_for(r->len), dim3(kBlock), 0, r->stream, static_cast<c32*>(r->psi),
                           static_cast<const c32*>(other->psi), r->len, coef<float>(c, 0), coef<float>(c, 1));
    QVNT_HIP(hipGetLastError());
    return order_after(other->stream, r->stream);
}

int qvnt_reg_set_num(qvnt_reg* r, uint32_t q_num) {
    if (!r) return fail(QVNT_E_ARG, "null register");
    if (!r->owns_memory) return fail(QVNT_E_ARG, "register memory is not its own");
    if (q_num > 40) return fail(QVNT_E_OOM, "register too large");
    const uint32_t ns = q_num < 3 ? 3 : q_num;
    const uint64_t len = 1ull << ns;
    const uint64_t keep = std::min<uint64_t>(r->len, 1ull << q_num);
    void* psi = nullptr;
    if (dev_alloc(r->device, &psi, len * r->amp_bytes()) != QVNT_OK)
        return fail(QVNT_E_OOM, "device memory for the resized state");
    if (r->prec == QVNT_F64)
        hipLaunchKernelGGL(k_resize<double>, grid_for(len), dim3(kBlock), 0, r->stream, static_cast<c64*>(psi), len,
                           static_cast<const c64*>(r->psi), keep);
    else
        hipLaunchKernelGGL(k_resize<float>, grid_for(len), dim3(kBlock), 0, r->stream, static_cast<c32*>(psi), len,
                           static_cast<const c32*>(r->psi), keep);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(r->stream);
        dev_release(r->device, psi, len * r->amp_bytes());
        return fail(QVNT_E_HIP, std::string("set_num: ") + hipGetErrorString(e));
    }
    dev_release(r->device, r->psi, r->len * r->amp_bytes());
    r->psi = psi;
    r->q_num = q_num;
    r->ns = ns;
    r->len = len;
    return QVNT_OK;
}

int qvnt_reg_get_polar(qvnt_reg* r, double* dst, uint64_t len) {
    if (!r) return fail(QVNT_E_ARG, "null register");
    const uint64_t n = 1ull << r->q_num;
    if (!dst || len < 2 * n) return fail(QVNT_E_ARG, "destination must hold 2 * 2^n doubles");
    const uint64_t chunk = std::min<uint64_t>(n, 1ull << 23);
    void* vbuf = nullptr;
    if (dev_alloc(r->device, &vbuf, 2 * chunk * sizeof(double)) != QVNT_OK) return fail(QVNT_E_OOM, "polar staging buffer");
    double* buf = static_cast<double*>(vbuf);
    int rc = QVNT_OK;
    for (uint64_t off = 0; off < n && rc == QVNT_OK; off += chunk) {
        const uint64_t m = std::min(chunk, n - off);
        if (r->prec == QVNT_F64)
            hipLaunchKernelGGL(k_polar<double>, grid_for(m), dim3(kBlock), 0, r->stream, static_cast<const c64*>(r->psi) + off,
                               m, buf);
        else
            hipLaunchKernelGGL(k_polar<float>, grid_for(m), dim3(kBlock), 0, r->stream, static_cast<const c32*>(r->psi) + off,
                               m, buf);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(dst + 2 * off, buf, 2 * m * sizeof(double), hipMemcpyDeviceToHost, r->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
        if (e != hipSuccess) rc = fail(QVNT_E_HIP, std::string("get_polar: ") + hipGetErrorString(e));
    }
    if (rc != QVNT_OK) (void)hipStreamSynchronize(r->stream);
    dev_release(r->device, buf, 2 * chunk * sizeof(double));
    return rc;
}

int qvnt_reg_sample_all(qvnt_reg* r, uint64_t shots, uint64_t seed, uint64_t stream, uint64_t* hist, uint64_t len) {
    if (!r) return fail(QVNT_E_ARG, "null register");
    if (!hist || len < r->len) return fail(QVNT_E_ARG, "histogram must hold one count per buffer element");
    std::fill(hist, hist + r->len, 0ull);
    const uint64_t chunk = std::min<uint64_t>(shots, 1ull << 20);
    std::vector<double> u(chunk);
    std::vector<uint64_t> outs(chunk);
    for (uint64_t k0 = 0; k0 < shots; k0 += chunk) {
        const uint64_t m = std::min(chunk, shots - k0);
        for (uint64_t k = 0; k < m; ++k) u[k] = qvnt_uniform(seed, stream, k0 + k);
        double total = 0.0;
        if (int rc = tree_sample(r, 1ull << r->q_num, u.data(), m, outs.data(), &total)) return rc;
        if (!(total > 0.0) || !std::isfinite(total)) return fail(QVNT_E_ZERO_NORM, "sampling a zero-norm or non-finite state");
        for (uint64_t k = 0; k < m; ++k) hist[outs[k]] += 1;
    }
    return QVNT_OK;
}

int qvnt_reg_run_program(qvnt_reg* r, const qvnt_step* steps, size_t n, const qvnt_op* tail, uint32_t measure_xor,
                         uint64_t* c_value, uint32_t c_num, const double* draws, size_t n_draws, size_t* used) {
    if (!r || !c_value || (n && !steps)) return fail(QVNT_E_ARG, "null argument");
    const uint64_t c_mask = c_num >= 64 ? ~0ull : ((1ull << c_num) - 1u);
    const uint64_t q_mask = (1ull << r->q_num) - 1u;
    size_t k_draw = 0;
    int rc = QVNT_OK;
    for (size_t k = 0; k < n && rc == QVNT_OK; ++k) {
        const qvnt_step& s = steps[k];
        if (s.sep > QVNT_SEP_RESET) {
            rc = fail(QVNT_E_ARG, "unknown separator in step " + std::to_string(k));
            break;
        }
        if (s.sep == QVNT_SEP_IF) {
            if (get_by_mask(*c_value, s.a & c_mask) == s.b && s.op) rc = qvnt_reg_apply(r, s.op);
            continue;
        }
        if (s.op) rc = qvnt_reg_apply(r, s.op);
        if (rc) break;
        if (s.sep == QVNT_SEP_MEASURE) {
            uint64_t outcome = 0;
            if (s.a & q_mask) {
                if (k_draw >= n_draws) {
                    rc = fail(QVNT_E_ARG, "program needs more draws than given");
                    break;
                }
                rc = qvnt_reg_measure_mask(r, s.a, draws[k_draw++], 0, &outcome);
                if (rc) break;
            }
            for (uint64_t qm = s.a, cm = s.b; qm && cm; qm &= qm - 1, cm &= cm - 1) {
                const uint64_t qb = qm & (~qm + 1), cb = cm & (~cm + 1);
                const bool bit = (outcome & qb) != 0;
                if (measure_xor) {
                    if (bit) *c_value ^= cb;
                } else if (bit) {
                    *c_value |= cb;
                } else {
                    *c_value &= ~cb;
                }
            }
        } else if (s.sep == QVNT_SEP_RESET) {
            rc = qvnt_reg_reset_by_mask(r, s.a);
        }
    }
    if (rc == QVNT_OK && tail) rc = qvnt_reg_apply(r, tail);
    if (used) *used = k_draw;
    return rc;
}

}  // extern "C"
