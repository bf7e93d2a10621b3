// The plain form of QReg::apply: one in-place kernel per SingleOp.
//
// The reference gathers every output element from a read-only input buffer
// (src/operator/atomic/dispatch.rs:32-47, multi/mod.rs:96-104), which needs
// a second state-sized buffer and streams the whole state even for
// controlled gates.  Here each thread owns whole index orbits of the gate
// (1, 2 or 4 amplitudes that the gate maps onto each other), reads them,
// and writes them back in place:
//   - only the control-satisfied subspace is enumerated: control bits are
//     deposited as fixed 1s into the compact thread index, so no lane is
//     wasted on a failed predicate;
//   - orbits that the gate leaves unchanged (even-parity pairs of the swap
//     family, the |0> half of single-bit phase gates) are not touched;
//   - consecutive threads get consecutive free index bits, so every wave
//     load is one or two contiguous runs whatever the target qubit.
// FP contraction is off in this file: outputs equal the reference's to the
// last bit in FP64.
#pragma clang fp contract(off)

#include <cstdlib>

#include "device.hpp"
#include "reg.hpp"

namespace qvnt {
namespace {

constexpr int kBlock = 256;       // k_init
constexpr int kMaxBlock = 1024;   // gate kernels: block size chosen at launch (block_setting)

// A dispatch carries its size in work-items as a 32-bit number, so grids
// are capped at kMaxGrid blocks and every kernel strides over chunks of
// blockDim * U orbits.
constexpr uint64_t kMaxGrid = 1u << 20;

// Loads and stores of one amplitude; NT marks them non-temporal (streamed
// through the caches with early eviction), used when no other access of
// the same launch touches the line.
template <bool NT, class R>
__device__ __forceinline__ cplx<R> ld(const cplx<R>* p) {
    if constexpr (NT) return cplx<R>{__builtin_nontemporal_load(&p->re), __builtin_nontemporal_load(&p->im)};
    else return *p;
}
template <bool NT, class R>
__device__ __forceinline__ void st(cplx<R>* p, cplx<R> v) {
    if constexpr (NT) {
        __builtin_nontemporal_store(v.re, &p->re);
        __builtin_nontemporal_store(v.im, &p->im);
    } else {
        *p = v;
    }
}

template <int K, class R, int U, bool NT>
__global__ __launch_bounds__(kMaxBlock) void k_diag(cplx<R>* __restrict__ psi, uint64_t groups, Deposit dep, Par g) {
  const uint64_t B = blockDim.x, per = B * U;
  const uint64_t nchunks = (groups + per - 1) / per;
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t base = chunk * per + threadIdx.x;
    cplx<R> v[U];
    uint64_t idx[U];
    bool in[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint64_t t = (U == 2 && dep.skew) ? unit_index(chunk * B + threadIdx.x, (uint32_t)u, dep.skew)
                                                : base + (uint64_t)u * B;
        idx[u] = deposit(t, dep);
        in[u] = t < groups;
        if (in[u]) v[u] = ld<NT>(psi + idx[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (in[u]) st<NT>(psi + idx[u], acts(g, idx[u]) ? f_diag<K>(g, idx[u], v[u]) : v[u]);
  }
}

// FP32 diagonals, two amplitudes per 16-byte access: a work item is the
// group indices 2t and 2t + 1, which the deposit maps to neighbouring
// amplitudes idx and idx + 1 when bit 0 is not a fixed (control) bit
// (launch_kernel checks that, and that the group count is even); idx is
// even, so the access is 16-byte aligned.  One 8-byte amplitude per lane
// left FP32 diagonals at 5.2 TB/s where FP64 ones (16 bytes per lane)
// reach 6.4; two per lane: 5.8 TB/s (32 qubits, TFIM ring).
typedef float f32x4_t __attribute__((ext_vector_type(4)));
template <int K, int U, bool NT>
__global__ __launch_bounds__(kMaxBlock) void k_diag_x2(cplx<float>* __restrict__ psi, uint64_t pairs, Deposit dep,
                                                       Par g) {
  const uint64_t B = blockDim.x, per = B * U;
  const uint64_t nchunks = (pairs + per - 1) / per;
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t base = chunk * per + threadIdx.x;
    f32x4_t q[U];
    uint64_t idx[U];
    bool in[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint64_t t = base + (uint64_t)u * B;
        idx[u] = deposit(2 * t, dep);
        in[u] = t < pairs;
        const f32x4_t* a = reinterpret_cast<const f32x4_t*>(psi + idx[u]);
        if (in[u]) q[u] = NT ? __builtin_nontemporal_load(a) : *a;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (!in[u]) continue;
        cplx<float> a{q[u].x, q[u].y}, b{q[u].z, q[u].w};
        if (acts(g, idx[u])) a = f_diag<K>(g, idx[u], a);
        if (acts(g, idx[u] + 1)) b = f_diag<K>(g, idx[u] + 1, b);
        const f32x4_t o = {a.re, a.im, b.re, b.im};
        f32x4_t* d = reinterpret_cast<f32x4_t*>(psi + idx[u]);
        if (NT) __builtin_nontemporal_store(o, d);
        else *d = o;
    }
  }
}

// FP32 pairs, two orbits per work item (2t and 2t + 1: neighbouring
// amplitudes i0, i0 + 1 and their partners, when bit 0 is neither fixed
// nor in the pair mask; launch_kernel checks that), so every access is 16
// bytes, as in k_diag_x2.
template <int K, int U, bool NT>
__global__ __launch_bounds__(kMaxBlock) void k_pair_x2(cplx<float>* __restrict__ psi, uint64_t pairs, Deposit dep,
                                                       Par g) {
  const uint64_t B = blockDim.x, per = B * U;
  const uint64_t nchunks = (pairs + per - 1) / per;
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t base = chunk * per + threadIdx.x;
    f32x4_t q0[U], q1[U];
    uint64_t i0[U];
    bool in[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint64_t t = base + (uint64_t)u * B;
        i0[u] = deposit(2 * t, dep);
        in[u] = t < pairs;
        if (in[u]) {
            const f32x4_t* a0 = reinterpret_cast<const f32x4_t*>(psi + i0[u]);
            const f32x4_t* a1 = reinterpret_cast<const f32x4_t*>(psi + (i0[u] ^ g.a));
            q0[u] = NT ? __builtin_nontemporal_load(a0) : *a0;
            q1[u] = NT ? __builtin_nontemporal_load(a1) : *a1;
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (!in[u]) continue;
        const uint64_t j0 = i0[u], j1 = i0[u] ^ g.a;
        cplx<float> x0{q0[u].x, q0[u].y}, y0{q0[u].z, q0[u].w};   // j0, j0 + 1
        cplx<float> x1{q1[u].x, q1[u].y}, y1{q1[u].z, q1[u].w};   // j1, j1 + 1
        cplx<float> ox0 = x0, ox1 = x1, oy0 = y0, oy1 = y1;
        if (acts(g, j0)) {
            ox0 = f_pair<K>(g, j0, x0, x1);
            ox1 = f_pair<K>(g, j1, x1, x0);
        }
        if (acts(g, j0 + 1)) {
            oy0 = f_pair<K>(g, j0 + 1, y0, y1);
            oy1 = f_pair<K>(g, j1 + 1, y1, y0);
        }
        const f32x4_t o0 = {ox0.re, ox0.im, oy0.re, oy0.im}, o1 = {ox1.re, ox1.im, oy1.re, oy1.im};
        f32x4_t* d0 = reinterpret_cast<f32x4_t*>(psi + j0);
        f32x4_t* d1 = reinterpret_cast<f32x4_t*>(psi + j1);
        if (NT) {
            __builtin_nontemporal_store(o0, d0);
            __builtin_nontemporal_store(o1, d1);
        } else {
            *d0 = o0;
            *d1 = o1;
        }
    }
  }
}

template <int K, class R, int U, bool NT>
__global__ __launch_bounds__(kMaxBlock) void k_pair(cplx<R>* __restrict__ psi, uint64_t groups, Deposit dep, Par g) {
  const uint64_t B = blockDim.x, per = B * U;
  const uint64_t nchunks = (groups + per - 1) / per;
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t base = chunk * per + threadIdx.x;
    cplx<R> v0[U], v1[U];
    uint64_t i0[U];
    bool in[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint64_t t = (U == 2 && dep.skew) ? unit_index(chunk * B + threadIdx.x, (uint32_t)u, dep.skew)
                                                : base + (uint64_t)u * B;
        i0[u] = deposit(t, dep);
        in[u] = t < groups;
        if (in[u]) {
            v0[u] = ld<NT>(psi + i0[u]);
            v1[u] = ld<NT>(psi + (i0[u] ^ g.a));
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (in[u]) {
            const uint64_t i1 = i0[u] ^ g.a;
            const bool on = acts(g, i0[u]);   // i0 and i1 differ only in the orbit mask
            const cplx<R> o0 = on ? f_pair<K>(g, i0[u], v0[u], v1[u]) : v0[u];
            const cplx<R> o1 = on ? f_pair<K>(g, i1, v1[u], v0[u]) : v1[u];
            st<NT>(psi + i0[u], o0);
            st<NT>(psi + i1, o1);
        }
    }
  }
}

template <int K, class R, int U, bool NT>
__global__ __launch_bounds__(kMaxBlock) void k_quad(cplx<R>* __restrict__ psi, uint64_t groups, Deposit dep, Par g) {
  const uint64_t B = blockDim.x, per = B * U;
  const uint64_t nchunks = (groups + per - 1) / per;
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t base = chunk * per + threadIdx.x;
    cplx<R> v[U][4];
    uint64_t i0[U];
    bool in[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint64_t t = (U == 2 && dep.skew) ? unit_index(chunk * B + threadIdx.x, (uint32_t)u, dep.skew)
                                                : base + (uint64_t)u * B;
        i0[u] = deposit(t, dep);
        in[u] = t < groups;
        if (in[u]) {
            v[u][0] = ld<NT>(psi + i0[u]);
            v[u][1] = ld<NT>(psi + (i0[u] | g.a));
            v[u][2] = ld<NT>(psi + (i0[u] | g.b));
            v[u][3] = ld<NT>(psi + (i0[u] | g.a | g.b));
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (in[u]) {
            cplx<R> o[4];
            const bool on = acts(g, i0[u]);
#pragma unroll
            for (int s = 0; s < 4; ++s) o[s] = on ? f_quad<K>(g, s, v[u]) : v[u][s];
            st<NT>(psi + i0[u], o[0]);
            st<NT>(psi + (i0[u] | g.a), o[1]);
            st<NT>(psi + (i0[u] | g.b), o[2]);
            st<NT>(psi + (i0[u] | g.a | g.b), o[3]);
        }
    }
  }
}

// Pair gates whose orbit mask lies in the lane bits (qubits 0..5, no
// control among them): each lane loads one amplitude (a wave reads one
// contiguous run), takes its partner's value from lane ^ a with a
// cross-lane shuffle, and writes its own output -- the formula of the
// reference for that amplitude (f_pair), so the result is the same to the
// bit.  groups = amplitudes of the control subspace, a multiple of 64, so
// every wave is either whole or idle and all of its lanes shuffle.
template <int K, class R, int U>
__global__ __launch_bounds__(kMaxBlock) void k_pair_lanes(cplx<R>* __restrict__ psi, uint64_t groups, Deposit dep,
                                                          Par g) {
    const uint64_t B = blockDim.x, per = B * U;
    const uint64_t nchunks = (groups + per - 1) / per;
    const int lm = (int)g.a;
    for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
        const uint64_t base = chunk * per + threadIdx.x;
        cplx<R> v[U];
        uint64_t idx[U];
        bool in[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            // units 2^skew apart (skew >= 6 keeps the lane bits below them)
            const uint64_t t = (U == 4 && dep.skew) ? unit_index(chunk * B + threadIdx.x, (uint32_t)u, dep.skew, 2)
                                                    : base + (uint64_t)u * B;
            idx[u] = deposit(t, dep);
            in[u] = t < groups;
            v[u] = cplx<R>{(R)0, (R)0};
            if (in[u]) v[u] = ld<true>(psi + idx[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const cplx<R> other{__shfl_xor(v[u].re, lm), __shfl_xor(v[u].im, lm)};
            if (in[u]) st<true>(psi + idx[u], acts(g, idx[u]) ? f_pair<K>(g, idx[u], v[u], other) : v[u]);
        }
    }
}

// H2 gates whose two bits lie in the lane bits (qubits 0..5, no deposited
// control among them): each lane loads one amplitude and takes the other
// three of its quad from lanes ^ a, ^ b, ^ (a|b); the output is the
// reference's formula for its slot (h2.rs:22-38, f_h2_rel), bit for bit.
template <class R, int U>
__global__ __launch_bounds__(kMaxBlock) void k_h2_lanes(cplx<R>* __restrict__ psi, uint64_t groups, Deposit dep, Par g) {
    const uint64_t B = blockDim.x, per = B * U;
    const uint64_t nchunks = (groups + per - 1) / per;
    const int la = (int)g.a, lb = (int)g.b;
    const int s = ((threadIdx.x & la) ? 1 : 0) | ((threadIdx.x & lb) ? 2 : 0);   // this lane's slot
    for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
        const uint64_t base = chunk * per + threadIdx.x;
        cplx<R> v[U];
        uint64_t idx[U];
        bool in[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t t = (U == 4 && dep.skew) ? unit_index(chunk * B + threadIdx.x, (uint32_t)u, dep.skew, 2)
                                                    : base + (uint64_t)u * B;
            idx[u] = deposit(t, dep);
            in[u] = t < groups;
            v[u] = cplx<R>{(R)0, (R)0};
            if (in[u]) v[u] = ld<true>(psi + idx[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const cplx<R> xa{__shfl_xor(v[u].re, la), __shfl_xor(v[u].im, la)};
            const cplx<R> xb{__shfl_xor(v[u].re, lb), __shfl_xor(v[u].im, lb)};
            const cplx<R> xab{__shfl_xor(v[u].re, la | lb), __shfl_xor(v[u].im, la | lb)};
            const cplx<R> o = down<R>(f_h2_rel(s, up(v[u]), up(xa), up(xb), up(xab)));
            if (in[u]) st<true>(psi + idx[u], acts(g, idx[u]) ? o : v[u]);
        }
    }
}

// H2 gates with one bit in the lane bits (qubits 0..5) and the other,
// m, above them: each lane loads amplitudes idx and idx | m (two contiguous
// runs per wave) and takes the lane partner of each from lane ^ l; the
// output is the reference's formula for its slot (f_h2_rel), bit for bit.
template <class R, int U>
__global__ __launch_bounds__(kMaxBlock) void k_h2_half(cplx<R>* __restrict__ psi, uint64_t groups, Deposit dep, Par g) {
    const uint64_t B = blockDim.x, per = B * U;
    const uint64_t nchunks = (groups + per - 1) / per;
    const bool lane_is_a = g.a < 64;                       // which of a, b lies in the lane bits
    const int l = (int)(lane_is_a ? g.a : g.b);
    const uint64_t m = lane_is_a ? g.b : g.a;
    const int sl = (threadIdx.x & l) ? 1 : 0;              // this lane's bit l
    for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
        const uint64_t base = chunk * per + threadIdx.x;
        cplx<R> v0[U], v1[U];
        uint64_t idx[U];
        bool in[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t t = (U == 2 && dep.skew) ? unit_index(chunk * B + threadIdx.x, (uint32_t)u, dep.skew)
                                                    : base + (uint64_t)u * B;
            idx[u] = deposit(t, dep);                      // bit m is 0
            in[u] = t < groups;
            v0[u] = v1[u] = cplx<R>{(R)0, (R)0};
            if (in[u]) {
                v0[u] = ld<true>(psi + idx[u]);
                v1[u] = ld<true>(psi + (idx[u] | m));
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const c64 p0 = up(v0[u]), p1 = up(v1[u]);
            const c64 q0 = up(cplx<R>{__shfl_xor(v0[u].re, l), __shfl_xor(v0[u].im, l)});
            const c64 q1 = up(cplx<R>{__shfl_xor(v1[u].re, l), __shfl_xor(v1[u].im, l)});
            // slots: bit 0 = a, bit 1 = b; from idx (bit m = 0) and from idx | m
            const int s0 = lane_is_a ? sl : (sl << 1), s1 = lane_is_a ? (sl | 2) : ((sl << 1) | 1);
            // t1 = value at ^a, t2 = at ^b, t3 = at ^ab
            const c64 o0 = lane_is_a ? f_h2_rel(s0, p0, q0, p1, q1) : f_h2_rel(s0, p0, p1, q0, q1);
            const c64 o1 = lane_is_a ? f_h2_rel(s1, p1, q1, p0, q0) : f_h2_rel(s1, p1, p0, q1, q0);
            if (in[u]) {
                const bool on = acts(g, idx[u]);
                st<true>(psi + idx[u], on ? down<R>(o0) : v0[u]);
                st<true>(psi + (idx[u] | m), on ? down<R>(o1) : v1[u]);
            }
        }
    }
}

// QReg::with_state (quant.rs:129-150): zeros, psi[state] = 1.
template <class R>
__global__ __launch_bounds__(kBlock) void k_init(cplx<R>* __restrict__ psi, uint64_t len, uint64_t state) {
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < len; i += (uint64_t)gridDim.x * kBlock)
        psi[i] = cplx<R>{i == state ? (R)1 : (R)0, (R)0};
}

// ----------------------------------------------------------- host side
// Orbits per thread and threads per block of the gate kernels (tuning
// knobs; the defaults are the measured best on a 2^30-amplitude state).
// QVNT_PAIR_LANES=0 keeps low-qubit pair gates on k_pair (for comparison).
bool lane_setting() {
    static const bool on = [] {
        const char* s = std::getenv("QVNT_PAIR_LANES");
        return !(s && std::atoi(s) == 0);
    }();
    return on;
}
// Orbits per thread: pair kernels two (placed 128 KiB apart, see
// skew_setting), quad kernels one (measured at 30 qubits: two skewed quads
// were 3 % faster on qubits 28-29 but 3-9 % slower on 12-13 and 2-20);
// QVNT_UNROLL_PAIR / QVNT_UNROLL_QUAD (1, 2 or 4) override.
int units_from_env(const char* name, int def) {
    const char* s = std::getenv(name);
    const int v = s ? std::atoi(s) : def;
    return (v == 1 || v == 2 || v == 4) ? v : def;
}
int pair_unroll_setting(GClass cls) {
    static const int pair = units_from_env("QVNT_UNROLL_PAIR", 2);
    static const int quad = units_from_env("QVNT_UNROLL_QUAD", 1);
    return cls == GClass::Pair ? pair : quad;
}
int unroll_setting() {
    static int u = [] {
        const char* s = std::getenv("QVNT_UNROLL");
        int v = s ? std::atoi(s) : 2;
        return (v == 1 || v == 2 || v == 4) ? v : 2;
    }();
    return u;
}
// Diagonal and pair kernels with two units per thread place them 2^S apart
// in the compact index: two streams 128 KiB apart.  Measured in FP64:
// k_diag gains 3-5 % from 28 to 31 qubits (6.08 -> 6.35 TB/s at 30), but
// above kSkewMaxBytes one stream is as fast or faster (34 qubits: 6.13
// against 5.62 TB/s), so diagonals of larger states keep one; k_pair gains
// at 30 qubits (5.71 -> 6.04 TB/s) and loses nothing at 34.
// Quad kernels place their two orbits the same way as pairs.
// QVNT_DIAG_SKEW / QVNT_PAIR_SKEW / QVNT_QUAD_SKEW override S at every size
// (0: the units blockDim apart).
constexpr uint64_t kSkewMaxBytes = 32ull << 30;
uint32_t skew_setting(GClass cls, size_t amp_bytes, uint64_t state_bytes) {
    static const int diag_env = [] {
        const char* s = std::getenv("QVNT_DIAG_SKEW");
        return s ? std::atoi(s) : -1;
    }();
    static const int pair_env = [] {
        const char* s = std::getenv("QVNT_PAIR_SKEW");
        return s ? std::atoi(s) : -1;
    }();
    static const int quad_env = [] {
        const char* s = std::getenv("QVNT_QUAD_SKEW");
        return s ? std::atoi(s) : -1;
    }();
    const int env = cls == GClass::Diag ? diag_env : cls == GClass::Pair ? pair_env : quad_env;
    if (env >= 0) return env <= 40 ? (uint32_t)env : 0u;
    if (cls == GClass::Diag && state_bytes > kSkewMaxBytes) return 0u;
    return amp_bytes == 16 ? 13u : 14u;  // 128 KiB
}

// The lane kernel's four units per lane 2^12 apart in the compact index
// (measured, ry on qubits 0-5: FP64 at 30 qubits 5.84 TB/s one block apart,
// 6.35 at 2^12, 5.97 at 2^13, 5.86 at 2^14; FP32 at 31 qubits 5.58, 6.57,
// 6.36, 5.91); like the diagonals, only up to kSkewMaxBytes of state.
// QVNT_LANES_SKEW overrides S (0: one block apart).
uint32_t lanes_skew_setting(size_t amp_bytes, uint64_t state_bytes) {
    static const int env = [] {
        const char* s = std::getenv("QVNT_LANES_SKEW");
        return s ? std::atoi(s) : -1;
    }();
    if (env >= 0) return env <= 40 ? (uint32_t)env : 0u;
    (void)amp_bytes;
    return state_bytes > kSkewMaxBytes ? 0u : 12u;
}

unsigned block_setting() {
    static unsigned b = [] {
        const char* s = std::getenv("QVNT_BLOCK");
        int v = s ? std::atoi(s) : 256;
        return (unsigned)((v == 256 || v == 512 || v == 1024) ? v : 256);
    }();
    return b;
}

void fix_bits(Deposit& d, uint64_t bits, bool one) {
    d.zeros |= bits;
    d.n = popc(d.zeros);
    if (one) d.set |= bits;
}

// Gate parameters in FP64 for both amplitude precisions; pred / podd: the
// index bits tested per amplitude (Par).
Par make_par(const qvnt_gate& g, uint64_t pred, uint64_t podd) {
    Par p{};
    p.pred = pred;
    p.podd = podd;
    p.a = g.a;
    p.b = g.b;
    p.re = g.re;
    p.im = g.im;
    for (int k = 0; k < 32; ++k) p.m[k] = g.m[k];
    p.dagger = g.dagger ? 1u : 0u;
    p.i_pow = (~(uint32_t)(popc(g.a) + 1u)) & 3u;  // atomic/y.rs:11
    return p;
}

template <class R, int U, bool NT>
int launch_kernel(qvnt_reg* r, const qvnt_gate& gate, GClass cls, uint64_t groups, const Deposit& d0, const Par& p) {
    Deposit d = d0;
    const uint32_t S = U == 2 ? skew_setting(cls, sizeof(cplx<R>), r->len * sizeof(cplx<R>)) : 0u;
    d.skew = (S && groups >= (2ull << S)) ? S : 0u;
    cplx<R>* psi = static_cast<cplx<R>*>(r->psi);
    const unsigned B = block_setting();
    const uint64_t nchunks = (groups + (uint64_t)B * U - 1) / ((uint64_t)B * U);
    const dim3 grid((unsigned)(nchunks < kMaxGrid ? nchunks : kMaxGrid));
    const dim3 block(B);
    hipStream_t s = r->stream;
#define QVNT_CASE(KIND, KERNEL)                                            \
    case KIND:                                                             \
        hipLaunchKernelGGL((KERNEL<KIND, R, U, NT>), grid, block, 0, s, psi, groups, d, p); \
        break;
    if constexpr (sizeof(R) == 4) {
        if (cls == GClass::Diag && (d.zeros & 1u) == 0 && (groups & 1u) == 0) {
            const uint64_t pairs = groups / 2;
            const uint64_t nx2 = (pairs + (uint64_t)B * U - 1) / ((uint64_t)B * U);
            const dim3 gx2((unsigned)(nx2 < kMaxGrid ? nx2 : kMaxGrid));
            cplx<float>* pf = reinterpret_cast<cplx<float>*>(psi);
#define QVNT_X2(KIND)                                                                        \
    case KIND:                                                                               \
        hipLaunchKernelGGL((k_diag_x2<KIND, U, NT>), gx2, block, 0, s, pf, pairs, d0, p);    \
        break;
            switch (gate.kind) {
                QVNT_X2(QVNT_Z)
                QVNT_X2(QVNT_S)
                QVNT_X2(QVNT_T)
                QVNT_X2(QVNT_RZ)
                QVNT_X2(QVNT_RZZ)
                QVNT_X2(QVNT_PHI)
                default: return fail(QVNT_E_ARG, "not a diagonal kind");
            }
#undef QVNT_X2
            QVNT_HIP(hipGetLastError());
            return QVNT_OK;
        }
        // (pairs only for states above kSkewMaxBytes: below, the skewed
        // one-amplitude kernel is faster -- TFIM ring FP32 at 32 qubits
        // (32 GiB): 86.3 gates/s against 83.8; at 35 qubits the two-amplitude
        // kernel wins, 11.35 against 11.07)
        if (cls == GClass::Pair && r->len * sizeof(cplx<R>) > kSkewMaxBytes && (d.zeros & 1u) == 0 && (p.a & 1u) == 0 &&
            (groups & 1u) == 0) {
            const uint64_t pairs = groups / 2;
            const uint64_t nx2 = (pairs + (uint64_t)B * U - 1) / ((uint64_t)B * U);
            const dim3 gx2((unsigned)(nx2 < kMaxGrid ? nx2 : kMaxGrid));
            cplx<float>* pf = reinterpret_cast<cplx<float>*>(psi);
#define QVNT_X2(KIND)                                                                        \
    case KIND:                                                                               \
        hipLaunchKernelGGL((k_pair_x2<KIND, U, NT>), gx2, block, 0, s, pf, pairs, d0, p);    \
        break;
            switch (gate.kind) {
                QVNT_X2(QVNT_X)
                QVNT_X2(QVNT_Y)
                QVNT_X2(QVNT_RX)
                QVNT_X2(QVNT_RXX)
                QVNT_X2(QVNT_RY)
                QVNT_X2(QVNT_RYY)
                QVNT_X2(QVNT_H1)
                QVNT_X2(QVNT_U1)
                QVNT_X2(QVNT_SWAP)
                QVNT_X2(QVNT_ISWAP)
                QVNT_X2(QVNT_SQRT_SWAP)
                QVNT_X2(QVNT_SQRT_ISWAP)
                default: return fail(QVNT_E_ARG, "not a pair kind");
            }
#undef QVNT_X2
            QVNT_HIP(hipGetLastError());
            return QVNT_OK;
        }
    }
    if (cls == GClass::Diag) {
        switch (gate.kind) {
            QVNT_CASE(QVNT_Z, k_diag)
            QVNT_CASE(QVNT_S, k_diag)
            QVNT_CASE(QVNT_T, k_diag)
            QVNT_CASE(QVNT_RZ, k_diag)
            QVNT_CASE(QVNT_RZZ, k_diag)
            QVNT_CASE(QVNT_PHI, k_diag)
            default: return fail(QVNT_E_ARG, "not a diagonal kind");
        }
    } else if (cls == GClass::Pair) {
        switch (gate.kind) {
            QVNT_CASE(QVNT_X, k_pair)
            QVNT_CASE(QVNT_Y, k_pair)
            QVNT_CASE(QVNT_RX, k_pair)
            QVNT_CASE(QVNT_RXX, k_pair)
            QVNT_CASE(QVNT_RY, k_pair)
            QVNT_CASE(QVNT_RYY, k_pair)
            QVNT_CASE(QVNT_H1, k_pair)
            QVNT_CASE(QVNT_U1, k_pair)
            QVNT_CASE(QVNT_SWAP, k_pair)
            QVNT_CASE(QVNT_ISWAP, k_pair)
            QVNT_CASE(QVNT_SQRT_SWAP, k_pair)
            QVNT_CASE(QVNT_SQRT_ISWAP, k_pair)
            default: return fail(QVNT_E_ARG, "not a pair kind");
        }
    } else {
        switch (gate.kind) {
            QVNT_CASE(QVNT_H2, k_quad)
            QVNT_CASE(QVNT_U2, k_quad)
            default: return fail(QVNT_E_ARG, "not a quad kind");
        }
    }
#undef QVNT_CASE
    QVNT_HIP(hipGetLastError());
    return QVNT_OK;
}

// Non-temporal accesses when every 128-byte line is read by one load
// instruction and written by one store (launch_unrolled); the lane kernels
// always qualify.
template <class R>
int launch_lanes(qvnt_reg* r, const qvnt_gate& gate, uint64_t pred) {
    Deposit d{};
    fix_bits(d, gate.ctrl & ~pred, true);
    const uint64_t groups = r->len >> d.n;
    const Par p = make_par(gate, pred, 0);
    cplx<R>* psi = static_cast<cplx<R>*>(r->psi);
    const unsigned B = block_setting();
    constexpr int U = 4;  // four amplitudes per lane: as many loads in flight as k_pair's two orbits
    const uint64_t nchunks = (groups + (uint64_t)B * U - 1) / ((uint64_t)B * U);
    // the four units of a lane 2^S apart, as the pair kernels' two orbits
    const uint32_t S = lanes_skew_setting(sizeof(cplx<R>), r->len * sizeof(cplx<R>));
    d.skew = (S >= 6 && groups >= (4ull << S)) ? S : 0u;
    const dim3 grid((unsigned)(nchunks < kMaxGrid ? nchunks : kMaxGrid)), block(B);
    hipStream_t s = r->stream;
#define QVNT_LANES(KIND)                                                                             \
    case KIND:                                                                                       \
        hipLaunchKernelGGL((k_pair_lanes<KIND, R, U>), grid, block, 0, s, psi, groups, d, p);       \
        break;
    switch (gate.kind) {
        QVNT_LANES(QVNT_X)
        QVNT_LANES(QVNT_Y)
        QVNT_LANES(QVNT_RX)
        QVNT_LANES(QVNT_RXX)
        QVNT_LANES(QVNT_RY)
        QVNT_LANES(QVNT_RYY)
        QVNT_LANES(QVNT_H1)
        QVNT_LANES(QVNT_U1)
        default: return fail(QVNT_E_ARG, "not a lane-pair kind");
    }
#undef QVNT_LANES
    QVNT_HIP(hipGetLastError());
    return QVNT_OK;
}

template <class R>
int launch_h2_lanes(qvnt_reg* r, const qvnt_gate& gate, uint64_t pred) {
    Deposit d{};
    fix_bits(d, gate.ctrl & ~pred, true);
    const uint64_t groups = r->len >> d.n;
    const Par p = make_par(gate, pred, 0);
    const unsigned B = block_setting();
    constexpr int U = 4;
    const uint64_t nchunks = (groups + (uint64_t)B * U - 1) / ((uint64_t)B * U);
    const uint32_t S = lanes_skew_setting(sizeof(cplx<R>), r->len * sizeof(cplx<R>));
    d.skew = (S >= 6 && groups >= (4ull << S)) ? S : 0u;
    const dim3 grid((unsigned)(nchunks < kMaxGrid ? nchunks : kMaxGrid)), block(B);
    hipLaunchKernelGGL((k_h2_lanes<R, U>), grid, block, 0, r->stream, static_cast<cplx<R>*>(r->psi), groups, d, p);
    QVNT_HIP(hipGetLastError());
    return QVNT_OK;
}

template <class R>
int launch_h2_half(qvnt_reg* r, const qvnt_gate& gate, uint64_t pred) {
    const uint64_t m = gate.a < 64 ? gate.b : gate.a;
    Deposit d{};
    fix_bits(d, m, false);
    fix_bits(d, gate.ctrl & ~pred, true);
    const uint64_t groups = r->len >> d.n;
    constexpr int U = 2;
    const uint32_t S = skew_setting(GClass::Pair, sizeof(cplx<R>), r->len * sizeof(cplx<R>));
    d.skew = (S >= 6 && groups >= (2ull << S)) ? S : 0u;
    const Par p = make_par(gate, pred, 0);
    const unsigned B = block_setting();
    const uint64_t nchunks = (groups + (uint64_t)B * U - 1) / ((uint64_t)B * U);
    const dim3 grid((unsigned)(nchunks < kMaxGrid ? nchunks : kMaxGrid)), block(B);
    hipLaunchKernelGGL((k_h2_half<R, U>), grid, block, 0, r->stream, static_cast<cplx<R>*>(r->psi), groups, d, p);
    QVNT_HIP(hipGetLastError());
    return QVNT_OK;
}

template <class R>
int launch_unrolled(qvnt_reg* r, const qvnt_gate& gate, GClass cls, uint64_t groups, const Deposit& d, const Par& p) {
    // every load and store instruction covers whole 128-byte lines exactly
    // when no index bit inside a line is fixed in the enumeration
    constexpr uint64_t line_amps = 128 / (2 * sizeof(R));
    const bool nt = (d.zeros & (line_amps - 1)) == 0;
    const int us = cls == GClass::Diag ? unroll_setting() : pair_unroll_setting(cls);
    if (us == 1)
        return nt ? launch_kernel<R, 1, true>(r, gate, cls, groups, d, p) : launch_kernel<R, 1, false>(r, gate, cls, groups, d, p);
    if (us == 4)
        return nt ? launch_kernel<R, 4, true>(r, gate, cls, groups, d, p) : launch_kernel<R, 4, false>(r, gate, cls, groups, d, p);
    return nt ? launch_kernel<R, 2, true>(r, gate, cls, groups, d, p) : launch_kernel<R, 2, false>(r, gate, cls, groups, d, p);
}

}  // namespace

int check_gate(uint64_t len, const qvnt_gate& gate, bool* noop) {
    const uint64_t lmask = len - 1u;
    const uint32_t k = gate.kind;
    *noop = false;
    if (k >= QVNT_KIND_COUNT) return fail(QVNT_E_ARG, "unknown gate kind");
    // Bit-count rules of the checked constructors (atomic/<gate>.rs is_valid).
    switch (k) {
    case QVNT_RX: case QVNT_RY: case QVNT_RZ: case QVNT_H1: case QVNT_U1: case QVNT_PHI:
        if (popc(gate.a) != 1) return fail(QVNT_E_INVALID_MASK, "Mask should contain 1 bit!");
        break;
    case QVNT_RXX: case QVNT_RYY: case QVNT_RZZ:
    case QVNT_SWAP: case QVNT_ISWAP: case QVNT_SQRT_SWAP: case QVNT_SQRT_ISWAP:
        if (popc(gate.a) != 2) return fail(QVNT_E_INVALID_MASK, "Mask should contain 2 bit!");
        break;
    case QVNT_H2: case QVNT_U2:
        if (popc(gate.a) != 1 || popc(gate.b) != 1 || gate.a == gate.b)
            return fail(QVNT_E_INVALID_MASK, "H2/U2 need two distinct single-bit masks");
        break;
    default: break;
    }
    const uint64_t touched = touched_bits(gate);
    if (gate.ctrl & touched) return fail(QVNT_E_CTRL_OVERLAP, "control mask overlaps the gate's qubits");
    // A control bit outside the buffer is never set: the predicate
    // `!idx & ctrl == 0` (dispatch.rs:35) fails everywhere, the gate is the
    // identity.
    if ((gate.ctrl & ~lmask) || k == QVNT_ID) {
        *noop = true;
        return QVNT_OK;
    }
    // Non-diagonal gates read psi[idx ^ mask]; beyond the buffer the
    // reference indexes out of bounds (a panic).  Diagonal gates only test
    // idx & mask, which is zero outside the buffer.
    if (!is_diagonal_kind(k) && (touched & ~lmask)) return fail(QVNT_E_ARG, "gate acts on qubits beyond the register");
    return QVNT_OK;
}

int launch_gate(qvnt_reg* r, const qvnt_gate& gate) {
    bool noop = false;
    if (int rc = check_gate(r->len, gate, &noop)) return rc;
    if (noop) return QVNT_OK;
    const uint64_t lmask = r->len - 1u;
    const uint32_t k = gate.kind;
    const bool diag = is_diagonal_kind(k);
    const bool quad = (k == QVNT_H2 || k == QVNT_U2);

    // Index bits inside one 128-byte line (8 FP64 or 16 FP32 amplitudes) are
    // never fixed in the enumeration: fixing one (a control, or the |1> half
    // of a one-qubit phase gate) would touch half of every line, and lines
    // written in part cost a read-modify-write in the memory system.  They
    // are tested per amplitude instead (Par::pred, Par::podd).
    const uint64_t line_mask = (r->prec == QVNT_F64 ? 8u : 16u) - 1u;
    const uint64_t pred = gate.ctrl & line_mask & lmask;
    uint64_t podd = 0;
    Deposit d{};
    GClass cls;
    const uint64_t a_in = gate.a & lmask;
    if (diag) {
        cls = GClass::Diag;
        const bool single_bit = popc(a_in) == 1;
        const bool zero_fixed = (k == QVNT_Z || k == QVNT_S || k == QVNT_T || k == QVNT_PHI);
        if (zero_fixed && a_in == 0) return QVNT_OK;               // identity
        // the |0> half is unchanged: skip it, unless the bit is inside a
        // line (the formula leaves it unchanged there)
        if (zero_fixed && single_bit && !(a_in & line_mask)) fix_bits(d, a_in, true);
    } else if (quad) {
        cls = GClass::Quad;
        fix_bits(d, gate.a | gate.b, false);
    } else {
        cls = GClass::Pair;
        if (gate.a == 0) return QVNT_OK;                           // X0 / Y0: identity
        const uint64_t hi = 1ull << msb(gate.a);
        fix_bits(d, hi, false);
        if (k == QVNT_SWAP || k == QVNT_ISWAP || k == QVNT_SQRT_SWAP || k == QVNT_SQRT_ISWAP) {
            // odd-parity orbits only (even ones are unchanged)
            const uint64_t lo = gate.a & ~hi;
            if (lo & line_mask) podd = gate.a;
            else fix_bits(d, lo, true);
        }
    }
    fix_bits(d, gate.ctrl & ~pred, true);
    const uint64_t groups = r->len >> d.n;

    const double bytes = 2.0 * (double)r->amp_bytes() * (double)(r->len >> popc(gate.ctrl));
    const int kc = cls == GClass::Diag ? QVNT_K_DIAG : cls == GClass::Pair ? QVNT_K_PAIR : QVNT_K_QUAD;
    Timed t(r, kc, bytes);
    // partners within a wave's 64 consecutive amplitudes: lane exchange
    const bool swap_family = k == QVNT_SWAP || k == QVNT_ISWAP || k == QVNT_SQRT_SWAP || k == QVNT_SQRT_ISWAP;
    const uint64_t hi_ctrl = gate.ctrl & ~pred;
    if (cls == GClass::Pair && !swap_family && (a_in & ~63ull) == 0 && (hi_ctrl & 63ull) == 0 &&
        r->ns >= 6 + popc(hi_ctrl & lmask) && lane_setting()) {
        if (r->prec == QVNT_F64) return launch_lanes<double>(r, gate, pred);
        return launch_lanes<float>(r, gate, pred);
    }
    const bool a_lane = gate.a < 64, b_lane = gate.b < 64;
    if (cls == GClass::Quad && k == QVNT_H2 && (a_lane || b_lane) && (hi_ctrl & 63ull) == 0 && lane_setting()) {
        if (a_lane && b_lane && r->ns >= 6 + popc(hi_ctrl & lmask)) {
            if (r->prec == QVNT_F64) return launch_h2_lanes<double>(r, gate, pred);
            return launch_h2_lanes<float>(r, gate, pred);
        }
        if (!(a_lane && b_lane) && r->ns >= 7 + popc(hi_ctrl & lmask)) {
            if (r->prec == QVNT_F64) return launch_h2_half<double>(r, gate, pred);
            return launch_h2_half<float>(r, gate, pred);
        }
    }
    const Par p = make_par(gate, pred, podd);
    if (r->prec == QVNT_F64) return launch_unrolled<double>(r, gate, cls, groups, d, p);
    return launch_unrolled<float>(r, gate, cls, groups, d, p);
}

int launch_init(qvnt_reg* r, uint64_t state) {
    const uint64_t nb = (r->len + kBlock - 1) / kBlock;
    const dim3 grid((unsigned)(nb < kMaxGrid ? nb : kMaxGrid));
    Timed t(r, QVNT_K_OTHER, (double)r->amp_bytes() * (double)r->len);
    if (r->prec == QVNT_F64)
        hipLaunchKernelGGL(k_init<double>, grid, dim3(kBlock), 0, r->stream, static_cast<c64*>(r->psi), r->len, state);
    else
        hipLaunchKernelGGL(k_init<float>, grid, dim3(kBlock), 0, r->stream, static_cast<c32*>(r->psi), r->len, state);
    QVNT_HIP(hipGetLastError());
    return QVNT_OK;
}

}  // namespace qvnt
