// Measurement: QReg::get_absolute / get_probabilities / measure_mask
// (src/register/quant.rs:434-507) on the device.
//
// The reference sums |psi|^2 sequentially, builds rand's WeightedIndex
// cumulative array on the host thread and draws from an unseeded thread_rng
// (quant.rs:437-442, 496-497).  Its draw cannot be reproduced, so a
// deterministic contract replaces it (the CPU oracle implements the same):
//   w_i = re*re + im*im in FP64 (no FMA), node(0,j) = w_j,
//   node(l,j) = node(l-1,2j) + node(l-1,2j+1)        (pairwise tree),
//   draw: x = u * node(top,0); from the root, with children (L, R), go left
//   iff x < L or R == 0, else x -= L and go right; the leaf reached is r.
// Tree levels from 2^11 leaves upwards are stored (O(2^n / 2^11) doubles);
// a draw recomputes the 2^11-leaf subtree it lands in.  FP32 registers are
// reduced, searched and rescaled in FP64 as well.
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "device.hpp"
#include "reg.hpp"

namespace qvnt {
namespace {

constexpr int kBlock = 256;
constexpr unsigned kLeafLog = 11;          // leaves per stored bottom node
constexpr unsigned kLeaf = 1u << kLeafLog;

template <class R>
__device__ __forceinline__ double weight(cplx<R> p) {
    const double re = (double)p.re, im = (double)p.im;
    return re * re + im * im;
}

// node(11, blockIdx.x): each wave reduces 512 consecutive leaves, 64 per
// load; lanes pair up with xor-shuffles (adjacent pairs first), then the 8
// chunk sums and the 4 wave sums pair up in the same adjacent order.
// A dispatch carries its size in work-items as a 32-bit number: grids are
// capped at kMaxGrid blocks and kernels stride over their work.
constexpr uint64_t kMaxGrid = 1u << 20;

template <class R>
__global__ __launch_bounds__(kBlock) void k_leaf(const cplx<R>* __restrict__ psi, uint64_t nleaves, double* __restrict__ out) {
  for (uint64_t blk = blockIdx.x; blk < nleaves; blk += gridDim.x) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t base = blk * kLeaf + (uint64_t)wave * 512u;
    double s[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] = weight(psi[base + (uint64_t)j * 64u + (uint64_t)lane]);
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
#pragma unroll
        for (int j = 0; j < 8; ++j) s[j] = s[j] + __shfl_xor(s[j], o);
    const double w = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
    __shared__ double sh[4];
    if (lane == 0) sh[wave] = w;
    __syncthreads();
    if (threadIdx.x == 0) out[blk] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
  }
}

// Pairwise levels of `span` (<= 2048, a power of two) values held in LDS:
// lv + off(k) with off(k) = 2*span - (2*span >> k).
__device__ void lds_levels(double* lv, unsigned span) {
    unsigned in_off = 0, m = span;
    while (m > 1) {
        const unsigned h = m >> 1, out_off = in_off + m;
        for (unsigned t = threadIdx.x; t < h; t += kBlock) lv[out_off + t] = lv[in_off + 2 * t] + lv[in_off + 2 * t + 1];
        __syncthreads();
        in_off = out_off;
        m = h;
    }
}

// Whole tree of a buffer of <= 2048 leaves: writes node(top, 0).
template <class R>
__global__ __launch_bounds__(kBlock) void k_small(const cplx<R>* __restrict__ psi, unsigned count, double* __restrict__ out) {
    __shared__ double lv[2 * kLeaf];
    for (unsigned t = threadIdx.x; t < count; t += kBlock) lv[t] = weight(psi[t]);
    __syncthreads();
    lds_levels(lv, count);
    if (threadIdx.x == 0) out[0] = lv[2 * count - 2];
}

struct LevelPtrs {
    double* p[kLeafLog + 1];
};

// Levels L+1 .. L+log2(span) above a stored level L of n_in nodes; every
// block owns `span` consecutive nodes.
__global__ __launch_bounds__(kBlock) void k_upper(const double* __restrict__ in, unsigned span, LevelPtrs out) {
    __shared__ double lv[2 * kLeaf];
    const uint64_t base = (uint64_t)blockIdx.x * span;
    for (unsigned t = threadIdx.x; t < span; t += kBlock) lv[t] = in[base + t];
    __syncthreads();
    lds_levels(lv, span);
    unsigned off = span, m = span >> 1;
    for (unsigned k = 1; m >= 1; ++k, off += m, m >>= 1)
        for (unsigned t = threadIdx.x; t < m; t += kBlock) out.p[k][(uint64_t)blockIdx.x * m + t] = lv[off + t];
}

struct TreeInfo {
    uint32_t top;      // log2(count)
    uint32_t lb;       // lowest stored level, min(top, 11)
    uint64_t off[64];  // offset of stored level l (lb <= l <= top) in the tree array
};

// One draw per block: descend the stored levels, then the recomputed
// 2^lb-leaf subtree of the block reached.  Draw k is u[k], or u0 when u is
// null (one draw, passed by value: no copy to the device before the launch).
template <class R>
__global__ __launch_bounds__(kBlock) void k_sample(const cplx<R>* __restrict__ psi, const double* __restrict__ tree,
                                                   TreeInfo info, const double* __restrict__ u, double u0,
                                                   uint64_t* __restrict__ out, double* __restrict__ total) {
    __shared__ double lv[2 * kLeaf];
    __shared__ uint64_t sj;
    __shared__ double sx;
    if (threadIdx.x == 0) {
        const double T = tree[info.off[info.top]];
        double x = (u ? u[blockIdx.x] : u0) * T;
        uint64_t j = 0;
        for (uint32_t l = info.top; l > info.lb; --l) {
            const double L = tree[info.off[l - 1] + 2 * j], Rr = tree[info.off[l - 1] + 2 * j + 1];
            if (x < L || Rr == 0.0) {
                j = 2 * j;
            } else {
                x = x - L;
                j = 2 * j + 1;
            }
        }
        sj = j;
        sx = x;
        if (blockIdx.x == 0) total[0] = T;
    }
    __syncthreads();
    const unsigned span = 1u << info.lb;
    const uint64_t base = sj << info.lb;
    for (unsigned t = threadIdx.x; t < span; t += kBlock) lv[t] = weight(psi[base + t]);
    __syncthreads();
    lds_levels(lv, span);
    if (threadIdx.x == 0) {
        double x = sx;
        uint64_t k = 0;
        // level l of the block subtree lives at off(l) = 2*span - (2*span >> l)
        for (uint32_t l = info.lb; l > 0; --l) {
            const unsigned off = 2 * span - ((2 * span) >> (l - 1));
            const double L = lv[off + 2 * k], Rr = lv[off + 2 * k + 1];
            if (x < L || Rr == 0.0) {
                k = 2 * k;
            } else {
                x = x - L;
                k = 2 * k + 1;
            }
        }
        out[blockIdx.x] = base + k;
    }
}

// collapse_mask (quant.rs:468-486): zero every amplitude with
// (idx ^ r) & mask != 0; store-only.
template <class R>
__global__ __launch_bounds__(kBlock) void k_collapse(cplx<R>* __restrict__ psi, uint64_t len, uint64_t r, uint64_t mask) {
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < len; i += (uint64_t)gridDim.x * kBlock)
        if ((i ^ r) & mask) psi[i] = cplx<R>{(R)0, (R)0};
}

// Renormalisation of the kept subspace (the bits of `mask` fixed to r's),
// multiplied in FP64.
template <class R>
__global__ __launch_bounds__(kBlock) void k_rescale(cplx<R>* __restrict__ psi, uint64_t groups, Deposit dep, double scale) {
    for (uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x; t < groups; t += (uint64_t)gridDim.x * kBlock) {
        const uint64_t i = deposit(t, dep);
        const cplx<R> p = psi[i];
        psi[i] = cplx<R>{(R)((double)p.re * scale), (R)((double)p.im * scale)};
    }
}

// get_probabilities (quant.rs:434-454): p_i = |psi_i|^2 * inv.
template <class R>
__global__ __launch_bounds__(kBlock) void k_probs(const cplx<R>* __restrict__ psi, uint64_t n, double inv, double* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) out[i] = weight(psi[i]) * inv;
}

// Comparison of two registers in FP64: per block, <a|b> (re, im), |a|^2,
// |b|^2 and max |a_i - b_i| over 4096 amplitudes; the host folds the
// partials in block order.
template <class RA, class RB>
__global__ __launch_bounds__(kBlock) void k_compare(const cplx<RA>* __restrict__ a, const cplx<RB>* __restrict__ b,
                                                    uint64_t len, double* __restrict__ part) {
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const uint64_t base = (uint64_t)blockIdx.x * 4096u;
    for (uint32_t k = threadIdx.x; k < 4096u; k += kBlock) {
        const uint64_t i = base + k;
        if (i >= len) break;
        const double ar = (double)a[i].re, ai = (double)a[i].im, br = (double)b[i].re, bi = (double)b[i].im;
        s[0] += ar * br + ai * bi;   // <a|b> = conj(a) b
        s[1] += ar * bi - ai * br;
        s[2] += ar * ar + ai * ai;
        s[3] += br * br + bi * bi;
        const double dr = ar - br, di = ai - bi;
        s[4] = fmax(s[4], sqrt(dr * dr + di * di));
    }
    __shared__ double sh[5][kBlock];
    for (int q = 0; q < 5; ++q) sh[q][threadIdx.x] = s[q];
    __syncthreads();
    for (unsigned h = kBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + h];
            sh[4][threadIdx.x] = fmax(sh[4][threadIdx.x], sh[4][threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) part[(uint64_t)blockIdx.x * 5 + threadIdx.x] = sh[threadIdx.x][0];
}

template <class RA, class RB>
void launch_compare(const void* a, const void* b, uint64_t len, double* part, unsigned blocks, hipStream_t s) {
    hipLaunchKernelGGL((k_compare<RA, RB>), dim3(blocks), dim3(kBlock), 0, s, static_cast<const cplx<RA>*>(a),
                       static_cast<const cplx<RB>*>(b), len, part);
}

unsigned log2u(uint64_t x) { return 63u - (unsigned)__builtin_clzll(x); }

TreeInfo tree_info(uint64_t count) {
    TreeInfo t{};
    t.top = log2u(count);
    t.lb = t.top < kLeafLog ? t.top : kLeafLog;
    uint64_t off = 0;
    for (uint32_t l = t.lb; l <= t.top; ++l) {
        t.off[l] = off;
        off += count >> l;
    }
    return t;
}

// Workspace blocks of a register that owns its memory come from and go
// back to the device's block cache (reg.hip); the register's stream is
// synchronized before a block goes back, so nothing queued still uses it.
int ws_alloc(qvnt_reg* r, void** p, size_t bytes) {
    if (r->owns_memory) return dev_alloc(r->device, p, bytes);
    if (hipMalloc(p, bytes) == hipSuccess) return QVNT_OK;
    (void)hipGetLastError();
    *p = nullptr;
    return QVNT_E_OOM;
}

int ws_release(qvnt_reg* r, void* p, size_t bytes) {
    if (!p) return QVNT_OK;
    if (!r->owns_memory) {
        QVNT_HIP(hipFree(p));
        return QVNT_OK;
    }
    QVNT_HIP(hipStreamSynchronize(r->stream));
    dev_release(r->device, p, bytes);
    return QVNT_OK;
}

int ensure_workspace(qvnt_reg* r, size_t tree_doubles, size_t shots) {
    if (tree_doubles > r->tree_cap) {
        void* old = r->tree;
        const size_t old_bytes = r->tree_cap * sizeof(double);
        r->tree = nullptr;
        r->tree_cap = 0;
        if (int rc = ws_release(r, old, old_bytes)) return rc;
        void* p = nullptr;
        if (ws_alloc(r, &p, tree_doubles * sizeof(double)) != QVNT_OK)
            return fail(QVNT_E_OOM, "device memory for the measurement tree");
        r->tree = static_cast<double*>(p);
        r->tree_cap = tree_doubles;
    }
    if (shots > r->shots_cap || !r->dscratch) {
        const size_t cap = shots < 64 ? 64 : shots;
        // (both blocks were sized by shots_cap; it is cleared before either
        // goes, so a failure below leaves no block recorded at a wrong size)
        void* d_old = r->dscratch;
        void* i_old = r->iscratch;
        const size_t old_cap = r->shots_cap;
        r->dscratch = nullptr;
        r->iscratch = nullptr;
        r->shots_cap = 0;
        if (int rc = ws_release(r, d_old, (old_cap + 1) * sizeof(double))) return rc;
        if (int rc = ws_release(r, i_old, (old_cap + 1) * sizeof(uint64_t))) return rc;
        void* d = nullptr;
        void* i = nullptr;
        if (ws_alloc(r, &d, (cap + 1) * sizeof(double)) != QVNT_OK)
            return fail(QVNT_E_OOM, "device memory for measurement scratch");
        if (ws_alloc(r, &i, (cap + 1) * sizeof(uint64_t)) != QVNT_OK) {
            (void)ws_release(r, d, (cap + 1) * sizeof(double));
            return fail(QVNT_E_OOM, "device memory for measurement scratch");
        }
        r->dscratch = static_cast<double*>(d);
        r->iscratch = static_cast<uint64_t*>(i);
        r->shots_cap = cap;
    }
    return QVNT_OK;
}

// The tree levels of the first `count` weights into `tree` (laid out by
// tree_info(count)); timed: bracket the launches with the register's timing
// events (the allocation-free path passes false: it creates no events).
template <class R>
int build_levels(qvnt_reg* r, double* tree, uint64_t count, const TreeInfo& info, bool timed) {
    const cplx<R>* psi = static_cast<const cplx<R>*>(r->psi);
    qvnt_reg* tr = timed ? r : nullptr;
    Timed t(tr, QVNT_K_REDUCE, (double)r->amp_bytes() * (double)count);
    if (info.top < kLeafLog) {
        hipLaunchKernelGGL(k_small<R>, dim3(1), dim3(kBlock), 0, r->stream, psi, (unsigned)count, tree + info.off[info.top]);
        QVNT_HIP(hipGetLastError());
        return QVNT_OK;
    }
    const uint64_t nleaves = count >> kLeafLog;
    hipLaunchKernelGGL(k_leaf<R>, dim3((unsigned)(nleaves < kMaxGrid ? nleaves : kMaxGrid)), dim3(kBlock), 0, r->stream, psi,
                       nleaves, tree + info.off[kLeafLog]);
    QVNT_HIP(hipGetLastError());
    for (uint32_t L = kLeafLog; L < info.top;) {
        const uint32_t steps = (info.top - L) < kLeafLog ? (info.top - L) : kLeafLog;
        const unsigned span = 1u << steps;
        const uint64_t n_in = count >> L;
        LevelPtrs lp{};
        for (uint32_t k = 1; k <= steps; ++k) lp.p[k] = tree + info.off[L + k];
        hipLaunchKernelGGL(k_upper, dim3((unsigned)(n_in / span)), dim3(kBlock), 0, r->stream,
                           (const double*)(tree + info.off[L]), span, lp);
        QVNT_HIP(hipGetLastError());
        L += steps;
    }
    return QVNT_OK;
}

// Tree levels plus one draw per u_dev[k] (u_dev null: the single draw u0)
// into out_dev, total into total_dev[0]; every pointer is device memory; no
// synchronisation.
int sample_into(qvnt_reg* r, double* tree, uint64_t count, const double* u_dev, double u0, size_t shots,
                uint64_t* out_dev, double* total_dev, bool timed) {
    const TreeInfo info = tree_info(count);
    int rc = r->prec == QVNT_F64 ? build_levels<double>(r, tree, count, info, timed)
                                 : build_levels<float>(r, tree, count, info, timed);
    if (rc) return rc;
    Timed t(timed ? r : nullptr, QVNT_K_OTHER, 0.0);
    if (r->prec == QVNT_F64)
        hipLaunchKernelGGL(k_sample<double>, dim3((unsigned)shots), dim3(kBlock), 0, r->stream,
                           static_cast<const c64*>(r->psi), (const double*)tree, info, u_dev, u0, out_dev, total_dev);
    else
        hipLaunchKernelGGL(k_sample<float>, dim3((unsigned)shots), dim3(kBlock), 0, r->stream,
                           static_cast<const c32*>(r->psi), (const double*)tree, info, u_dev, u0, out_dev, total_dev);
    QVNT_HIP(hipGetLastError());
    return QVNT_OK;
}

}  // namespace

int tree_build(qvnt_reg* r, uint64_t count, double* total_out) {
    const TreeInfo info = tree_info(count);
    int rc = ensure_workspace(r, (size_t)(info.off[info.top] + 1), 1);
    if (rc) return rc;
    rc = r->prec == QVNT_F64 ? build_levels<double>(r, r->tree, count, info, true)
                             : build_levels<float>(r, r->tree, count, info, true);
    if (rc) return rc;
    if (total_out) {
        QVNT_HIP(hipMemcpyAsync(total_out, r->tree + info.off[info.top], sizeof(double), hipMemcpyDeviceToHost, r->stream));
        QVNT_HIP(hipStreamSynchronize(r->stream));
    }
    return QVNT_OK;
}

size_t tree_doubles(uint64_t count) {
    const TreeInfo info = tree_info(count);
    return (size_t)(info.off[info.top] + 1);
}

int sample_launch(qvnt_reg* r, uint64_t count, const double* u_dev, size_t shots, uint64_t* out_dev, double* total_dev) {
    if (tree_doubles(count) > r->tree_cap) return fail(QVNT_E_ARG, "measurement tree workspace too small");
    return sample_into(r, r->tree, count, u_dev, 0.0, shots, out_dev, total_dev, true);
}

// The outcomes land in iscratch[0, shots) and the total right after them
// (iscratch holds shots_cap + 1 words), so one copy brings both back; a
// single draw goes to the kernel by value.
int tree_sample(qvnt_reg* r, uint64_t count, const double* u, size_t shots, uint64_t* out, double* total_out) {
    if (shots == 0) return fail(QVNT_E_ARG, "no draws");
    int rc = ensure_workspace(r, tree_doubles(count), shots);
    if (rc) return rc;
    if (tree_doubles(count) > r->tree_cap) return fail(QVNT_E_ARG, "measurement tree workspace too small");
    const double* u_dev = nullptr;
    if (shots > 1) {
        QVNT_HIP(hipMemcpyAsync(r->dscratch + 1, u, shots * sizeof(double), hipMemcpyHostToDevice, r->stream));
        u_dev = r->dscratch + 1;
    }
    double* total_dev = reinterpret_cast<double*>(r->iscratch + shots);
    rc = sample_into(r, r->tree, count, u_dev, shots > 1 ? 0.0 : u[0], shots, r->iscratch, total_dev, true);
    if (rc) return rc;
    uint64_t one[2];
    std::vector<uint64_t> many(shots > 1 ? shots + 1 : 0);
    uint64_t* host = shots > 1 ? many.data() : one;
    QVNT_HIP(hipMemcpyAsync(host, r->iscratch, (shots + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, r->stream));
    QVNT_HIP(hipStreamSynchronize(r->stream));
    std::memcpy(out, host, shots * sizeof(uint64_t));
    std::memcpy(total_out, host + shots, sizeof(double));
    return QVNT_OK;
}

int collapse(qvnt_reg* r, uint64_t outcome, uint64_t mask, bool renorm) {
    {
        const uint64_t nb = (r->len + kBlock - 1) / kBlock;
        const dim3 grid((unsigned)(nb < kMaxGrid ? nb : kMaxGrid));
        Timed t(r, QVNT_K_COLLAPSE, 2.0 * (double)r->amp_bytes() * (double)r->len);
        if (r->prec == QVNT_F64)
            hipLaunchKernelGGL(k_collapse<double>, grid, dim3(kBlock), 0, r->stream, static_cast<c64*>(r->psi), r->len, outcome, mask);
        else
            hipLaunchKernelGGL(k_collapse<float>, grid, dim3(kBlock), 0, r->stream, static_cast<c32*>(r->psi), r->len, outcome, mask);
        QVNT_HIP(hipGetLastError());
    }
    if (!renorm) return QVNT_OK;
    double kept = 0.0;
    int rc = tree_build(r, r->len, &kept);
    if (rc) return rc;
    if (!(kept > 0.0) || !std::isfinite(kept)) return fail(QVNT_E_ZERO_NORM, "collapsed state has zero norm");
    const double scale = 1.0 / std::sqrt(kept);
    Deposit d{};
    d.zeros = mask;
    d.n = popc(mask);
    d.set = outcome & mask;
    const uint64_t groups = r->len >> d.n;
    const uint64_t nb = (groups + kBlock - 1) / kBlock;
    const dim3 grid((unsigned)(nb < kMaxGrid ? nb : kMaxGrid));
    Timed t(r, QVNT_K_COLLAPSE, 2.0 * (double)r->amp_bytes() * (double)groups);
    if (r->prec == QVNT_F64)
        hipLaunchKernelGGL(k_rescale<double>, grid, dim3(kBlock), 0, r->stream, static_cast<c64*>(r->psi), groups, d, scale);
    else
        hipLaunchKernelGGL(k_rescale<float>, grid, dim3(kBlock), 0, r->stream, static_cast<c32*>(r->psi), groups, d, scale);
    QVNT_HIP(hipGetLastError());
    return QVNT_OK;
}

int rescale(qvnt_reg* r, double scale) {
    const Deposit d{};  // no fixed bits: every amplitude of the buffer
    const uint64_t nb = (r->len + kBlock - 1) / kBlock;
    const dim3 grid((unsigned)(nb < kMaxGrid ? nb : kMaxGrid));
    Timed t(r, QVNT_K_COLLAPSE, 2.0 * (double)r->amp_bytes() * (double)r->len);
    if (r->prec == QVNT_F64)
        hipLaunchKernelGGL(k_rescale<double>, grid, dim3(kBlock), 0, r->stream, static_cast<c64*>(r->psi), r->len, d, scale);
    else
        hipLaunchKernelGGL(k_rescale<float>, grid, dim3(kBlock), 0, r->stream, static_cast<c32*>(r->psi), r->len, d, scale);
    QVNT_HIP(hipGetLastError());
    return QVNT_OK;
}

int probabilities(qvnt_reg* r, double inv_total, double* dst, uint64_t n) {
    // bounded device staging: chunks of up to 2^24 probabilities
    const uint64_t chunk = n < (1ull << 24) ? n : (1ull << 24);
    void* vbuf = nullptr;
    if (dev_alloc(r->device, &vbuf, chunk * sizeof(double)) != QVNT_OK)
        return fail(QVNT_E_OOM, "probability staging buffer");
    double* buf = static_cast<double*>(vbuf);
    int rc = QVNT_OK;
    for (uint64_t off = 0; off < n && rc == QVNT_OK; off += chunk) {
        const uint64_t m = (n - off) < chunk ? (n - off) : chunk;
        const dim3 grid((unsigned)((m + kBlock - 1) / kBlock));
        if (r->prec == QVNT_F64)
            hipLaunchKernelGGL(k_probs<double>, grid, dim3(kBlock), 0, r->stream, static_cast<const c64*>(r->psi) + off, m, inv_total, buf);
        else
            hipLaunchKernelGGL(k_probs<float>, grid, dim3(kBlock), 0, r->stream, static_cast<const c32*>(r->psi) + off, m, inv_total, buf);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(dst + off, buf, m * sizeof(double), hipMemcpyDeviceToHost, r->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
        if (e != hipSuccess) rc = fail(QVNT_E_HIP, std::string("probabilities: ") + hipGetErrorString(e));
    }
    // (every chunk ended with a synchronization; after a failure, wait here)
    if (rc != QVNT_OK) (void)hipStreamSynchronize(r->stream);
    dev_release(r->device, buf, chunk * sizeof(double));
    return rc;
}

int compare(qvnt_reg* a, qvnt_reg* b, double* out) {
    const uint64_t len = a->len;
    const unsigned blocks = (unsigned)((len + 4095) / 4096);
    QVNT_HIP(hipStreamSynchronize(b->stream));
    void* vpart = nullptr;
    if (dev_alloc(a->device, &vpart, (size_t)blocks * 5 * sizeof(double)) != QVNT_OK)
        return fail(QVNT_E_OOM, "comparison partials");
    double* part = static_cast<double*>(vpart);
    if (a->prec == QVNT_F64 && b->prec == QVNT_F64) launch_compare<double, double>(a->psi, b->psi, len, part, blocks, a->stream);
    else if (a->prec == QVNT_F64) launch_compare<double, float>(a->psi, b->psi, len, part, blocks, a->stream);
    else if (b->prec == QVNT_F64) launch_compare<float, double>(a->psi, b->psi, len, part, blocks, a->stream);
    else launch_compare<float, float>(a->psi, b->psi, len, part, blocks, a->stream);
    std::vector<double> host((size_t)blocks * 5);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host.data(), part, host.size() * sizeof(double), hipMemcpyDeviceToHost, a->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(a->stream);
    if (e != hipSuccess) (void)hipStreamSynchronize(a->stream);
    dev_release(a->device, part, (size_t)blocks * 5 * sizeof(double));
    if (e != hipSuccess) return fail(QVNT_E_HIP, std::string("compare: ") + hipGetErrorString(e));
    double s[5] = {0, 0, 0, 0, 0};
    for (unsigned k = 0; k < blocks; ++k) {
        for (int q = 0; q < 4; ++q) s[q] += host[(size_t)k * 5 + q];
        s[4] = std::max(s[4], host[(size_t)k * 5 + 4]);
    }
    for (int q = 0; q < 5; ++q) out[q] = s[q];
    return QVNT_OK;
}

}  // namespace qvnt

// ------------------------------------------------ allocation-free shots
// qvnt_reg_sample_mask_noalloc: the shots of qvnt_reg_sample_mask with the
// tree levels, draws, outcomes and total in `scratch`, a file-scope device
// array of kScratchBytes that every GPU the library runs on holds until the
// process ends.  In host code `scratch` names a host-side stand-in, not a
// device address: the device address comes from hipGetSymbolAddress, taken
// once per device with that device current and kept in g_scratch.  Every
// call on one GPU shares that GPU's array, so calls that overlap on one GPU
// race.  Nothing in it needs to start at zero: each call writes the levels,
// the draws, every outcome and the total before reading them.
namespace qvnt {
namespace {

constexpr size_t kScratchBytes = size_t(16) << 20;  // 16 MiB: registers up to 30 qubits
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

int scratch_address(int device, unsigned char** out) {
    if (device < 0 || device >= kMaxDevices) return fail(QVNT_E_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        int cur = 0;
        QVNT_HIP(hipGetDevice(&cur));
        if (cur != device) QVNT_HIP(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(QVNT_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[device] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch[device]);
    return QVNT_OK;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace
}  // namespace qvnt

extern "C" int qvnt_reg_sample_mask_noalloc(qvnt_reg* r, uint64_t mask, const double* u, size_t shots, uint64_t* out) {
    using namespace qvnt;
    if (!r) return fail(QVNT_E_ARG, "null register");
    if (shots == 0) return QVNT_OK;
    if (!u || !out) return fail(QVNT_E_ARG, "null draws or outcomes");
    if (shots > (1u << 20)) return fail(QVNT_E_ARG, "at most 2^20 shots per call");
    mask &= (1ull << r->q_num) - 1u;
    if (mask == 0) {
        for (size_t k = 0; k < shots; ++k) out[k] = 0;
        return QVNT_OK;
    }
    for (size_t k = 0; k < shots; ++k)
        if (!(u[k] >= 0.0 && u[k] < 1.0)) return fail(QVNT_E_ARG, "draws must lie in [0, 1)");
    // workspace layout: [tree levels | total | draws | outcomes], 256-byte aligned
    const uint64_t count = 1ull << r->q_num;
    const size_t off_total = align256(tree_doubles(count) * sizeof(double));
    const size_t off_u = off_total + 256;
    const size_t off_out = off_u + align256(shots * sizeof(double));
    const size_t need = off_out + align256(shots * sizeof(uint64_t));
    if (need > kScratchBytes)
        return fail(QVNT_E_ARG, "allocation-free shots need " + std::to_string(need) + " bytes of workspace; the "
                                "fixed workspace holds " + std::to_string(kScratchBytes) +
                                " (registers up to 30 qubits, fewer shots at 30)");
    unsigned char* ws = nullptr;
    if (int rc = scratch_address(r->device, &ws)) return rc;
    double* tree = reinterpret_cast<double*>(ws);
    double* total_dev = reinterpret_cast<double*>(ws + off_total);
    double* u_dev = reinterpret_cast<double*>(ws + off_u);
    uint64_t* out_dev = reinterpret_cast<uint64_t*>(ws + off_out);
    QVNT_HIP(hipMemcpyAsync(u_dev, u, shots * sizeof(double), hipMemcpyHostToDevice, r->stream));
    if (int rc = sample_into(r, tree, count, u_dev, 0.0, shots, out_dev, total_dev, false)) return rc;
    double total = 0.0;
    QVNT_HIP(hipMemcpyAsync(out, out_dev, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, r->stream));
    QVNT_HIP(hipMemcpyAsync(&total, total_dev, sizeof(double), hipMemcpyDeviceToHost, r->stream));
    QVNT_HIP(hipStreamSynchronize(r->stream));
    if (!(total > 0.0) || !std::isfinite(total)) return fail(QVNT_E_ZERO_NORM, "measurement of a zero-norm or non-finite state");
    for (size_t k = 0; k < shots; ++k) out[k] &= mask;
    return QVNT_OK;
}
