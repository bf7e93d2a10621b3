// C ABI of the device register: QReg::new / with_state / apply /
// get_absolute / get_probabilities / measure_mask (src/register/quant.rs),
// plus state copy-in/out and per-kernel-class timing.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "reg.hpp"

namespace qvnt {

// reg may be null: no timing (the allocation-free path creates no events).
Timed::Timed(qvnt_reg* reg, int kc, double by) : r(reg), kclass(kc), bytes(by) {
    if (!r || !r->timing) return;
    for (hipEvent_t* e : {&a, &b}) {
        if (!r->event_pool.empty()) {
            *e = r->event_pool.back();
            r->event_pool.pop_back();
        } else if (hipEventCreate(e) != hipSuccess) {
            *e = nullptr;
        }
    }
    if (a) (void)hipEventRecord(a, r->stream);
}

Timed::~Timed() {
    if (!r || !r->timing || !a || !b) return;
    (void)hipEventRecord(b, r->stream);
    r->pending.push_back({kclass, a, b, bytes});
}

namespace {

constexpr int kMaxDevices = 64;
constexpr size_t kStreamPool = 64;

struct DeviceCache {
    std::mutex mu;
    std::multimap<size_t, void*> blocks;  // by size
    size_t total = 0;
    std::vector<hipStream_t> streams;     // idle
};

// Never destroyed: a register freed while the process exits (after this
// library's static destructors, from a host's own static objects) still
// finds its device's cache.
DeviceCache& cache_of(int device) {
    static DeviceCache* const caches = new DeviceCache[kMaxDevices];
    return caches[device];
}

bool cached_device(int device) { return device >= 0 && device < kMaxDevices; }

void flush_blocks(int device) {
    if (!cached_device(device)) return;
    DeviceCache& c = cache_of(device);
    std::lock_guard<std::mutex> lk(c.mu);
    for (auto& kv : c.blocks) (void)hipFree(kv.second);
    c.blocks.clear();
    c.total = 0;
}

// Makes `device` current for its lifetime, so that a new block or stream is
// the device's own whatever device the calling thread has current.
struct OnDevice {
    int prev = -1;
    explicit OnDevice(int device) {
        int cur = 0;
        if (hipGetDevice(&cur) == hipSuccess && cur != device && hipSetDevice(device) == hipSuccess) prev = cur;
    }
    ~OnDevice() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

}  // namespace

int dev_alloc(int device, void** p, size_t bytes) {
    *p = nullptr;
    if (cached_device(device) && bytes <= kCacheBlock) {
        DeviceCache& c = cache_of(device);
        std::lock_guard<std::mutex> lk(c.mu);
        auto it = c.blocks.find(bytes);
        if (it != c.blocks.end()) {
            *p = it->second;
            c.total -= bytes;
            c.blocks.erase(it);
            return QVNT_OK;
        }
    }
    OnDevice on(device);
    if (hipMalloc(p, bytes) == hipSuccess) return QVNT_OK;
    (void)hipGetLastError();
    flush_blocks(device);
    if (hipMalloc(p, bytes) == hipSuccess) return QVNT_OK;
    (void)hipGetLastError();
    *p = nullptr;
    return QVNT_E_OOM;
}

void dev_release(int device, void* p, size_t bytes) {
    if (!p) return;
    if (cached_device(device) && bytes <= kCacheBlock) {
        DeviceCache& c = cache_of(device);
        std::lock_guard<std::mutex> lk(c.mu);
        if (c.total + bytes <= kCacheTotal) {
            c.blocks.emplace(bytes, p);
            c.total += bytes;
            return;
        }
    }
    (void)hipFree(p);
}

int stream_acquire(int device, hipStream_t* s) {
    *s = nullptr;
    if (cached_device(device)) {
        DeviceCache& c = cache_of(device);
        std::lock_guard<std::mutex> lk(c.mu);
        if (!c.streams.empty()) {
            *s = c.streams.back();
            c.streams.pop_back();
            return QVNT_OK;
        }
    }
    OnDevice on(device);
    if (hipStreamCreateWithFlags(s, hipStreamNonBlocking) == hipSuccess) return QVNT_OK;
    (void)hipGetLastError();
    *s = nullptr;
    return QVNT_E_HIP;
}

void stream_release(int device, hipStream_t s) {
    if (!s) return;
    if (cached_device(device)) {
        DeviceCache& c = cache_of(device);
        std::lock_guard<std::mutex> lk(c.mu);
        if (c.streams.size() < kStreamPool) {
            c.streams.push_back(s);
            return;
        }
    }
    (void)hipStreamDestroy(s);
}

namespace {

int fold_timing(qvnt_reg* r) {
    if (r->pending.empty()) return QVNT_OK;
    QVNT_HIP(hipStreamSynchronize(r->stream));
    for (auto& p : r->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            r->stats[p.kclass].launches += 1;
            r->stats[p.kclass].ms += ms;
            r->stats[p.kclass].bytes += p.bytes;
        }
        r->event_pool.push_back(p.a);
        r->event_pool.push_back(p.b);
    }
    r->pending.clear();
    return QVNT_OK;
}

// Only gfx950 code objects are built; refuse any other device up front.
// The device count and each device's architecture are read once per
// process (the count query alone cost ~65 us a register).
int check_device(int* dev) {
    static const int count = [] {
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess) {
            (void)hipGetLastError();
            c = 0;
        }
        return c;
    }();
    if (count == 0) return fail(QVNT_E_NO_DEVICE, "no HIP device visible");
    QVNT_HIP(hipGetDevice(dev));
    static std::atomic<bool> gfx950[kMaxDevices];
    if (cached_device(*dev) && gfx950[*dev].load(std::memory_order_acquire)) return QVNT_OK;
    hipDeviceProp_t prop;
    QVNT_HIP(hipGetDeviceProperties(&prop, *dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(QVNT_E_NO_DEVICE, std::string("device is ") + prop.gcnArchName + ", library is built for gfx950");
    if (cached_device(*dev)) gfx950[*dev].store(true, std::memory_order_release);
    return QVNT_OK;
}

int check_reg(const qvnt_reg* r) { return r ? QVNT_OK : fail(QVNT_E_ARG, "null register"); }

}  // namespace
}  // namespace qvnt

using namespace qvnt;

extern "C" {

int qvnt_cache_trim(void) {
    for (int d = 0; d < kMaxDevices; ++d) {
        flush_blocks(d);
        DeviceCache& c = cache_of(d);
        std::lock_guard<std::mutex> lk(c.mu);
        for (hipStream_t s : c.streams) (void)hipStreamDestroy(s);
        c.streams.clear();
    }
    return QVNT_OK;
}

int qvnt_reg_with_state(uint32_t q_num, uint64_t state, qvnt_prec prec, qvnt_reg** out) {
    if (!out) return fail(QVNT_E_ARG, "null output pointer");
    if (prec != QVNT_F64 && prec != QVNT_F32) return fail(QVNT_E_ARG, "unknown precision");
    if (q_num > 40) return fail(QVNT_E_OOM, "register too large");
    int dev = 0;
    int rc = check_device(&dev);
    if (rc) return rc;
    qvnt_reg* r = new (std::nothrow) qvnt_reg;
    if (!r) return fail(QVNT_E_OOM, "host allocation failed");
    r->q_num = q_num;
    r->ns = q_num < 3 ? 3 : q_num;
    r->len = 1ull << r->ns;
    r->prec = prec;
    r->device = dev;
    if (stream_acquire(dev, &r->stream) != QVNT_OK) {
        delete r;
        return fail(QVNT_E_HIP, "stream creation failed");
    }
    if (dev_alloc(dev, &r->psi, r->len * r->amp_bytes()) != QVNT_OK) {
        stream_release(dev, r->stream);
        delete r;
        return fail(QVNT_E_OOM, "device memory for the state vector");
    }
    const uint64_t q_mask = (1ull << q_num) - 1u;
    rc = launch_init(r, state & q_mask);
    if (rc) {
        qvnt_reg_free(r);
        return rc;
    }
    *out = r;
    clear_error();
    return QVNT_OK;
}

int qvnt_reg_new(uint32_t q_num, qvnt_prec prec, qvnt_reg** out) { return qvnt_reg_with_state(q_num, 0, prec, out); }

int qvnt_reg_set_dense(qvnt_reg* r, uint32_t min_ops) {
    if (!r) return fail(QVNT_E_ARG, "null register");
    r->dense_min = min_ops;
    return QVNT_OK;
}

int qvnt_reg_free(qvnt_reg* r) {
    if (!r) return QVNT_OK;
    (void)hipStreamSynchronize(r->stream);
    for (auto& p : r->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (hipEvent_t e : r->event_pool) (void)hipEventDestroy(e);
    if (r->owns_memory) {
        // (the stream is synchronized: nothing queued uses these any more)
        dev_release(r->device, r->psi, r->len * r->amp_bytes());
        dev_release(r->device, r->tree, r->tree_cap * sizeof(double));
        dev_release(r->device, r->dscratch, (r->shots_cap + 1) * sizeof(double));
        dev_release(r->device, r->iscratch, (r->shots_cap + 1) * sizeof(uint64_t));
        dev_release(r->device, r->prog, r->prog_cap);
        stream_release(r->device, r->stream);
    }
    delete r;
    return QVNT_OK;
}

uint32_t qvnt_reg_num(const qvnt_reg* r) { return r ? r->q_num : 0; }
uint64_t qvnt_reg_len(const qvnt_reg* r) { return r ? r->len : 0; }
qvnt_prec qvnt_reg_prec(const qvnt_reg* r) { return r ? r->prec : QVNT_F64; }

int qvnt_reg_apply_gates(qvnt_reg* r, const qvnt_gate* gates, size_t count) {
    if (int rc = check_reg(r)) return rc;
    if (count && !gates) return fail(QVNT_E_ARG, "null gate array");
    // validate the whole queue first: nothing is applied if a gate is bad
    for (size_t i = 0; i < count; ++i) {
        bool noop = false;
        if (int rc = check_gate(r->len, gates[i], &noop)) return rc;
    }
    for (size_t i = 0; i < count; ++i) {
        int rc = launch_gate(r, gates[i]);
        if (rc) return rc;
    }
    return QVNT_OK;
}

int qvnt_reg_apply_fused(qvnt_reg* r, const qvnt_op* op, uint32_t max_qubits, uint32_t* passes) {
    if (int rc = check_reg(r)) return rc;
    if (!op) return fail(QVNT_E_ARG, "null op");
    for (const qvnt_gate& g : op->g) {
        bool noop = false;
        if (int rc = check_gate(r->len, g, &noop)) return rc;
    }
    return apply_fused(r, op->g.data(), op->g.size(), max_qubits, passes);
}

int qvnt_fused_plan_stats(const qvnt_op* op, uint32_t q_num, qvnt_prec prec, uint32_t max_qubits, uint32_t dense_min,
                          double* out) {
    if (!op || !out) return fail(QVNT_E_ARG, "null op or output");
    if (q_num == 0 || q_num > 40) return fail(QVNT_E_ARG, "q_num must be 1..40");
    if (prec != QVNT_F64 && prec != QVNT_F32) return fail(QVNT_E_ARG, "unknown precision");
    const unsigned ns = q_num < 3 ? 3 : q_num;
    for (const qvnt_gate& g : op->g) {
        bool noop = false;
        if (int rc = check_gate(1ull << ns, g, &noop)) return rc;
    }
    fused_plan_stats(op->g.data(), op->g.size(), ns, prec == QVNT_F64, max_qubits, dense_min, out);
    return QVNT_OK;
}

int qvnt_reg_apply(qvnt_reg* r, const qvnt_op* op) {
    if (!op) return fail(QVNT_E_ARG, "null op");
    return qvnt_reg_apply_gates(r, op->g.data(), op->g.size());
}

int qvnt_reg_get_absolute(qvnt_reg* r, double* out) {
    if (int rc = check_reg(r)) return rc;
    if (!out) return fail(QVNT_E_ARG, "null output");
    return tree_build(r, r->len, out);
}

int qvnt_reg_get_probabilities(qvnt_reg* r, double* dst, uint64_t len) {
    if (int rc = check_reg(r)) return rc;
    const uint64_t n = 1ull << r->q_num;
    if (!dst || len < n) return fail(QVNT_E_ARG, "destination must hold 2^n doubles");
    double total = 0.0;
    int rc = tree_build(r, r->len, &total);
    if (rc) return rc;
    return probabilities(r, 1.0 / total, dst, n);
}

int qvnt_reg_sample_mask(qvnt_reg* r, uint64_t mask, const double* u, size_t shots, uint64_t* out) {
    if (int rc = check_reg(r)) return rc;
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
    double total = 0.0;
    int rc = tree_sample(r, 1ull << r->q_num, u, shots, out, &total);
    if (rc) return rc;
    if (!(total > 0.0) || !std::isfinite(total)) return fail(QVNT_E_ZERO_NORM, "measurement of a zero-norm or non-finite state");
    for (size_t k = 0; k < shots; ++k) out[k] &= mask;
    return QVNT_OK;
}

int qvnt_reg_measure_mask(qvnt_reg* r, uint64_t mask, double u, uint32_t flags, uint64_t* out) {
    if (int rc = check_reg(r)) return rc;
    if (!out) return fail(QVNT_E_ARG, "null output");
    mask &= (1ull << r->q_num) - 1u;   // quant.rs:491
    if (mask == 0) {                   // quant.rs:492-494: no draw, no collapse
        *out = 0;
        return QVNT_OK;
    }
    if (!(u >= 0.0 && u < 1.0)) return fail(QVNT_E_ARG, "draw must lie in [0, 1)");
    uint64_t outcome = 0;
    double total = 0.0;
    int rc = tree_sample(r, 1ull << r->q_num, &u, 1, &outcome, &total);
    if (rc) return rc;
    if (!(total > 0.0) || !std::isfinite(total)) return fail(QVNT_E_ZERO_NORM, "measurement of a zero-norm or non-finite state");
    rc = collapse(r, outcome, mask, (flags & QVNT_MEASURE_RENORM) != 0);
    if (rc) return rc;
    *out = outcome & mask;             // quant.rs:500
    return QVNT_OK;
}

int qvnt_reg_measure(qvnt_reg* r, double u, uint32_t flags, uint64_t* out) {
    if (int rc = check_reg(r)) return rc;
    return qvnt_reg_measure_mask(r, (1ull << r->q_num) - 1u, u, flags, out);
}

int qvnt_reg_compare(qvnt_reg* a, qvnt_reg* b, double out[5]) {
    if (int rc = check_reg(a)) return rc;
    if (int rc = check_reg(b)) return rc;
    if (!out) return fail(QVNT_E_ARG, "null output");
    if (a->len != b->len) return fail(QVNT_E_ARG, "registers differ in size");
    return compare(a, b, out);
}

int qvnt_reg_get_state(qvnt_reg* r, void* dst, size_t bytes) {
    if (int rc = check_reg(r)) return rc;
    if (!dst || bytes != r->len * r->amp_bytes()) return fail(QVNT_E_ARG, "state copy needs len * amplitude bytes");
    QVNT_HIP(hipMemcpyAsync(dst, r->psi, bytes, hipMemcpyDeviceToHost, r->stream));
    QVNT_HIP(hipStreamSynchronize(r->stream));
    return QVNT_OK;
}

int qvnt_reg_set_state(qvnt_reg* r, const void* src, size_t bytes) {
    if (int rc = check_reg(r)) return rc;
    if (!src || bytes != r->len * r->amp_bytes()) return fail(QVNT_E_ARG, "state copy needs len * amplitude bytes");
    QVNT_HIP(hipMemcpyAsync(r->psi, src, bytes, hipMemcpyHostToDevice, r->stream));
    QVNT_HIP(hipStreamSynchronize(r->stream));
    return QVNT_OK;
}

int qvnt_reg_debug(qvnt_reg* r, char* buf, size_t cap, size_t* need) {
    if (int rc = check_reg(r)) return rc;
    // Debug of QReg (quant.rs:603-623): the first min(2^n, 8) amplitudes
    // (MAX_LEN_TO_DISPLAY, quant.rs:16) as struct fields, ".." after them
    // when the register holds more
    constexpr uint64_t kMaxShown = 8;
    const uint64_t n = 1ull << r->q_num;
    const uint64_t shown = n < kMaxShown ? n : kMaxShown;
    double amps[2 * kMaxShown];
    if (r->prec == QVNT_F64) {
        QVNT_HIP(hipMemcpyAsync(amps, r->psi, shown * 16, hipMemcpyDeviceToHost, r->stream));
        QVNT_HIP(hipStreamSynchronize(r->stream));
    } else {
        float f[2 * kMaxShown];
        QVNT_HIP(hipMemcpyAsync(f, r->psi, shown * 8, hipMemcpyDeviceToHost, r->stream));
        QVNT_HIP(hipStreamSynchronize(r->stream));
        for (uint64_t k = 0; k < 2 * shown; ++k) amps[k] = (double)f[k];
    }
    std::string s = "QReg { ";
    for (uint64_t k = 0; k < shown; ++k)
        s += (k ? ", " : "") + std::to_string(k) + ": " + complex_debug_value(amps[2 * k], amps[2 * k + 1]);
    s += n > kMaxShown ? ", .. }" : " }";
    if (need) *need = s.size() + 1;
    if (buf && cap) {
        const size_t m = std::min(cap - 1, s.size());
        std::memcpy(buf, s.data(), m);
        buf[m] = '\0';
    }
    return QVNT_OK;
}

int qvnt_reg_sync(qvnt_reg* r) {
    if (int rc = check_reg(r)) return rc;
    QVNT_HIP(hipStreamSynchronize(r->stream));
    return QVNT_OK;
}

int qvnt_reg_set_timing(qvnt_reg* r, int enable) {
    if (int rc = check_reg(r)) return rc;
    if (int rc = fold_timing(r)) return rc;
    r->timing = enable != 0;
    return QVNT_OK;
}

int qvnt_reg_kernel_stats(qvnt_reg* r, qvnt_kstats* out, int reset) {
    if (int rc = check_reg(r)) return rc;
    if (int rc = fold_timing(r)) return rc;
    if (out) std::memcpy(out, r->stats, sizeof(r->stats));
    if (reset) std::memset(r->stats, 0, sizeof(r->stats));
    return QVNT_OK;
}

}  // extern "C"
