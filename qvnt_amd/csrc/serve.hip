// Continuous-batching request server (no reference counterpart: the
// reference runs one register at a time on the host).
//
// A request is one circuit on a fresh register followed by a batch of
// non-destructive shots of measure_mask.  The server owns
//   - one HBM arena (the admission budget), sub-allocated first-fit;
//   - `slots` HIP streams, each with a completion event, pinned host
//     staging for draws, fused programs and outcomes, and a worker thread.
// Requests queue at submit (validated, fused programs compiled on the
// caller's thread).  A slot's worker admits the first queued request that
// fits the arena -- smaller requests may pass a large one at the head of
// the queue, up to kMaxBypass times -- and enqueues its whole life on the
// slot's stream: init, apply (fused or plain), measurement tree and draws,
// outcome copy, event.  It then waits for the event, returns the arena
// block and posts the result.  Requests of 2^27 amplitudes or more run one
// at a time (kBigQubits); smaller ones overlap with them and each other.  Workers enqueue in parallel, so a long
// request's thousands of launches do not hold up admission into the other
// slots, and requests on different slots overlap on the device.  Kernels
// are deterministic (fixed reduction orders, no atomics), so a request's
// outcomes do not depend on what runs beside it.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_set>
#include <vector>

#include "fusion.hpp"
#include "reg.hpp"

namespace qvnt {
namespace {

constexpr size_t kAlign = 256;
constexpr int kMaxBypass = 8;
// Requests of at least 2^kBigQubits amplitudes run one at a time: each of
// their passes has tens of thousands of tiles, far more than the GPU holds
// at once, so two of them side by side only stretch each other (and the
// tail latency), while small requests still overlap with them.
constexpr unsigned kBigQubits = 27;
size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// First-fit sub-allocator over one device allocation.
class Arena {
public:
    int init(size_t bytes) {
        size_ = bytes & ~(kAlign - 1);
        if (hipMalloc(&base_, size_) != hipSuccess) {
            (void)hipGetLastError();
            return fail(QVNT_E_OOM, "device memory for the server arena");
        }
        free_[0] = size_;
        return QVNT_OK;
    }
    void destroy() {
        if (base_) (void)hipFree(base_);
        base_ = nullptr;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = align_up(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes;
            const size_t next = it->first + bytes;
            free_.erase(it);
            if (rest) free_[next] = rest;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = align_up(bytes);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    char* base() const { return static_cast<char*>(base_); }
    size_t size() const { return size_; }
    size_t peak() const { return peak_; }

private:
    void* base_ = nullptr;
    size_t size_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;
};

struct Record {
    qvnt_request req{};
    std::vector<qvnt_gate> gates;
    std::vector<fusion::Step> steps;
    std::vector<double> draws;
    uint64_t len = 0;
    size_t bytes = 0;  // arena bytes: state, tree, draws, outcomes, total, programs
    size_t prog_bytes = 0;
    bool big = false;  // q_num >= kBigQubits
    int bypassed = 0;
    qvnt_result res{};
    size_t arena_off = 0;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    unsigned char* pinned = nullptr;
    size_t pinned_cap = 0;
    qvnt_reg reg;
    Record* rec = nullptr;
    size_t out_off = 0, total_off = 0;  // offsets of outcomes / total in the pinned buffer
};

}  // namespace
}  // namespace qvnt

using namespace qvnt;

struct qvnt_server {
    Arena arena;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Record*> pending;
    std::deque<Record*> done;
    std::unordered_set<uint64_t> ids;  // ids submitted and not yet returned by poll
    int big_running = 0;               // admitted requests of >= kBigQubits
    std::vector<std::thread> workers;  // one per slot
    bool stop = false;
    int device = 0;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

    int start(Record* rec, Slot& sl);
    void finish(Slot& sl, int status);
    Record* admit();
    void work(Slot& sl);
};

// Runs without the server lock: touches only the slot, the record and the
// record's arena block (allocated by the caller).
int qvnt_server::start(Record* rec, Slot& sl) {
    const qvnt_request& q = rec->req;
    const size_t amp = q.prec == QVNT_F64 ? 16 : 8;
    // device layout of the arena block
    char* base = arena.base() + rec->arena_off;
    size_t off = 0;
    void* psi = base + off;
    off += align_up(rec->len * amp);
    double* tree = reinterpret_cast<double*>(base + off);
    const size_t tdbl = tree_doubles(1ull << q.q_num);
    off += align_up(tdbl * sizeof(double));
    double* draws = reinterpret_cast<double*>(base + off);
    off += align_up(q.shots * sizeof(double));
    uint64_t* outs = reinterpret_cast<uint64_t*>(base + off);
    off += align_up(q.shots * sizeof(uint64_t));
    double* total = reinterpret_cast<double*>(base + off);
    off += kAlign;
    char* progs = base + off;
    // pinned staging: draws, programs, outcomes, total
    const size_t need = align_up(q.shots * 8) + rec->prog_bytes + align_up(q.shots * 8) + kAlign;
    if (need > sl.pinned_cap) {
        if (sl.pinned) (void)hipHostFree(sl.pinned);
        sl.pinned = nullptr;
        sl.pinned_cap = 0;
        if (hipHostMalloc(reinterpret_cast<void**>(&sl.pinned), need, hipHostMallocDefault) != hipSuccess)
            return fail(QVNT_E_OOM, "pinned staging for a request");
        sl.pinned_cap = need;
    }
    // the register on arena memory and the slot's stream
    qvnt_reg& r = sl.reg;
    r.q_num = q.q_num;
    r.ns = q.q_num < 3 ? 3 : q.q_num;
    r.len = rec->len;
    r.prec = (qvnt_prec)q.prec;
    r.device = device;
    r.psi = psi;
    r.stream = sl.stream;
    r.tree = tree;
    r.tree_cap = tdbl;
    r.owns_memory = false;
    r.timing = false;

    int rc = launch_init(&r, q.init_state & ((1ull << q.q_num) - 1u));
    size_t poff = align_up(q.shots * 8), doff = 0;
    uint32_t passes = 0;
    for (size_t k = 0; rc == QVNT_OK && k < rec->steps.size(); ++k) {
        const fusion::Step& st = rec->steps[k];
        if (st.plain) {
            rc = launch_gate(&r, st.gate);
            ++passes;
            continue;
        }
        const size_t pb = program_bytes(st.prog);
        pack_program(st.prog, sl.pinned + poff);
        if (hipMemcpyAsync(progs + doff, sl.pinned + poff, pb, hipMemcpyHostToDevice, sl.stream) != hipSuccess)
            rc = fail(QVNT_E_HIP, "program upload");
        if (rc == QVNT_OK) rc = launch_program(&r, st.prog, progs + doff);
        passes += (uint32_t)st.prog.passes.size();
        poff += align_up(pb);
        doff += align_up(pb);
    }
    if (rc == QVNT_OK) {
        std::memcpy(sl.pinned, rec->draws.data(), q.shots * sizeof(double));
        if (hipMemcpyAsync(draws, sl.pinned, q.shots * sizeof(double), hipMemcpyHostToDevice, sl.stream) != hipSuccess)
            rc = fail(QVNT_E_HIP, "draw upload");
    }
    if (rc == QVNT_OK) rc = sample_launch(&r, 1ull << q.q_num, draws, q.shots, outs, total);
    sl.out_off = poff;
    sl.total_off = poff + align_up(q.shots * 8);
    if (rc == QVNT_OK &&
        (hipMemcpyAsync(sl.pinned + sl.out_off, outs, q.shots * sizeof(uint64_t), hipMemcpyDeviceToHost, sl.stream) != hipSuccess ||
         hipMemcpyAsync(sl.pinned + sl.total_off, total, sizeof(double), hipMemcpyDeviceToHost, sl.stream) != hipSuccess ||
         hipEventRecord(sl.done, sl.stream) != hipSuccess))
        rc = fail(QVNT_E_HIP, "result copy");
    rec->res.passes = passes;
    return rc;
}

void qvnt_server::finish(Slot& sl, int status) {
    Record* rec = sl.rec;
    const qvnt_request& q = rec->req;
    rec->res.t_done = now();
    rec->res.status = status;
    if (status == QVNT_OK) {
        double total = 0.0;
        std::memcpy(&total, sl.pinned + sl.total_off, sizeof(double));
        rec->res.total = total;
        if (!(total > 0.0)) rec->res.status = QVNT_E_ZERO_NORM;
        const uint64_t mask = q.mask & ((1ull << q.q_num) - 1u);
        const uint64_t* o = reinterpret_cast<const uint64_t*>(sl.pinned + sl.out_off);
        for (uint64_t k = 0; k < q.shots; ++k) q.outcomes[k] = o[k] & mask;
    }
    arena.release(rec->arena_off, rec->bytes);
    if (rec->big) --big_running;
    sl.rec = nullptr;
    done.push_back(rec);
}

// The first queued request that fits the arena (its block allocated), or
// null; called with the lock held.  The head of the queue may be passed
// over at most kMaxBypass times, then admission waits for space for it.
Record* qvnt_server::admit() {
    for (auto it = pending.begin(); it != pending.end(); ++it) {
        Record* rec = *it;
        size_t off = 0;
        if (rec->big && big_running > 0) continue;   // waits for the running large request
        if (!arena.alloc(rec->bytes, &off)) {
            if (it == pending.begin() && rec->bypassed >= kMaxBypass) return nullptr;
            continue;
        }
        if (it != pending.begin()) pending.front()->bypassed++;
        rec->arena_off = off;
        rec->res.t_admit = now();
        if (rec->big) ++big_running;
        pending.erase(it);
        return rec;
    }
    return nullptr;
}

// One slot's worker: admit, enqueue the request's whole life on the slot's
// stream (without the lock), wait for its completion event, post the result.
void qvnt_server::work(Slot& sl) {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        Record* rec = nullptr;
        cv_work.wait(lk, [&] { return stop || (!pending.empty() && (rec = admit()) != nullptr); });
        if (!rec) {
            if (stop) break;
            continue;
        }
        sl.rec = rec;
        lk.unlock();
        int rc = start(rec, sl);
        if (rc == QVNT_OK) {
            hipError_t e;
            while ((e = hipEventQuery(sl.done)) == hipErrorNotReady) std::this_thread::sleep_for(std::chrono::microseconds(20));
            if (e != hipSuccess) rc = QVNT_E_HIP;
        } else {
            (void)hipStreamSynchronize(sl.stream);
        }
        lk.lock();
        finish(sl, rc);
        cv_done.notify_all();
        cv_work.notify_all();   // arena space is free again
    }
}

extern "C" {

int qvnt_server_new(uint64_t hbm_bytes, uint32_t slots, qvnt_server** out) {
    if (!out || slots == 0 || slots > 64) return fail(QVNT_E_ARG, "need 1..64 slots and an output pointer");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(QVNT_E_NO_DEVICE, "no HIP device visible");
    hipDeviceProp_t prop;
    int dev = 0;
    QVNT_HIP(hipGetDevice(&dev));
    QVNT_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return fail(QVNT_E_NO_DEVICE, "library is built for gfx950");
    qvnt_server* s = new (std::nothrow) qvnt_server;
    if (!s) return fail(QVNT_E_OOM, "host allocation failed");
    s->device = dev;
    if (int rc = s->arena.init(hbm_bytes)) {
        delete s;
        return rc;
    }
    s->slots.resize(slots);
    for (Slot& sl : s->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
            qvnt_server_free(s);
            return fail(QVNT_E_HIP, "stream or event creation failed");
        }
    }
    for (Slot& sl : s->slots) s->workers.emplace_back([s, &sl] { s->work(sl); });
    *out = s;
    return QVNT_OK;
}

int qvnt_server_submit(qvnt_server* s, const qvnt_request* q) {
    if (!s || !q || !q->op) return fail(QVNT_E_ARG, "null server, request or operator");
    const double t_in = s->now();  // latency includes validation and the fused-program compile
    if (q->prec != QVNT_F64 && q->prec != QVNT_F32) return fail(QVNT_E_ARG, "unknown precision");
    if (q->q_num > 40) return fail(QVNT_E_OOM, "register too large");
    if (q->shots == 0 || q->shots > (1u << 20) || !q->draws || !q->outcomes)
        return fail(QVNT_E_ARG, "need 1..2^20 shots with draws and an outcome buffer");
    Record* rec = new (std::nothrow) Record;
    if (!rec) return fail(QVNT_E_OOM, "host allocation failed");
    rec->req = *q;
    rec->len = 1ull << (q->q_num < 3 ? 3 : q->q_num);
    rec->big = q->q_num >= kBigQubits;
    rec->gates = q->op->g;
    for (const qvnt_gate& g : rec->gates) {
        bool noop = false;
        if (int rc = check_gate(rec->len, g, &noop)) {
            delete rec;
            return rc;
        }
    }
    rec->draws.assign(q->draws, q->draws + q->shots);
    for (double u : rec->draws)
        if (!(u >= 0.0 && u < 1.0)) {
            delete rec;
            return fail(QVNT_E_ARG, "draws must lie in [0, 1)");
        }
    const unsigned ns = q->q_num < 3 ? 3 : q->q_num;
    if (q->fused && ns >= (unsigned)fusion::kMinFused) {
        // the same program qvnt_reg_apply_fused would run
        compile_fused(rec->gates.data(), rec->gates.size(), ns, q->prec == QVNT_F64, q->max_qubits, q->dense_min,
                      rec->steps);
    } else {
        for (const qvnt_gate& g : rec->gates) {
            fusion::Step st;
            st.plain = true;
            st.gate = g;
            rec->steps.push_back(st);
        }
    }
    for (const fusion::Step& st : rec->steps)
        if (!st.plain) rec->prog_bytes += align_up(program_bytes(st.prog));
    const size_t amp = q->prec == QVNT_F64 ? 16 : 8;
    rec->bytes = align_up(rec->len * amp) + align_up(tree_doubles(1ull << q->q_num) * 8) + 2 * align_up(q->shots * 8) +
                 kAlign + rec->prog_bytes;
    rec->res.id = q->id;
    std::lock_guard<std::mutex> lk(s->mu);
    // the caller's outcome buffer is written when the request finishes, and
    // poll returns results by id: a second request under a live id is refused
    if (!s->ids.insert(q->id).second) {
        delete rec;
        return fail(QVNT_E_ARG, "request id " + std::to_string(q->id) + " is already in flight");
    }
    rec->res.t_submit = t_in;
    if (rec->bytes > s->arena.size()) {
        rec->res.status = QVNT_E_OOM;
        rec->res.t_admit = rec->res.t_done = rec->res.t_submit;
        s->done.push_back(rec);
        s->cv_done.notify_all();
        return QVNT_OK;
    }
    s->pending.push_back(rec);
    s->cv_work.notify_all();
    return QVNT_OK;
}

int qvnt_server_poll(qvnt_server* s, qvnt_result* out, uint32_t cap, uint32_t* n, int wait_ms) {
    if (!s || !n || (cap && !out)) return fail(QVNT_E_ARG, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && wait_ms > 0) s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [s] { return !s->done.empty(); });
    uint32_t k = 0;
    while (k < cap && !s->done.empty()) {
        Record* rec = s->done.front();
        s->done.pop_front();
        out[k++] = rec->res;
        s->ids.erase(rec->res.id);
        delete rec;
    }
    *n = k;
    return QVNT_OK;
}

int qvnt_server_pending(qvnt_server* s, uint32_t* queued, uint32_t* running, uint64_t* peak_bytes) {
    if (!s) return fail(QVNT_E_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    uint32_t run = 0;
    for (const Slot& sl : s->slots) run += sl.rec ? 1u : 0u;
    if (queued) *queued = (uint32_t)s->pending.size();
    if (running) *running = run;
    if (peak_bytes) *peak_bytes = s->arena.peak();
    return QVNT_OK;
}

double qvnt_server_now(qvnt_server* s) { return s ? s->now() : 0.0; }

int qvnt_server_free(qvnt_server* s) {
    if (!s) return QVNT_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (Record* rec : s->pending) delete rec;
        s->pending.clear();
    }
    s->cv_work.notify_all();
    for (std::thread& w : s->workers)
        if (w.joinable()) w.join();
    for (Record* rec : s->done) delete rec;
    for (Slot& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.pinned) (void)hipHostFree(sl.pinned);
    }
    s->arena.destroy();
    delete s;
    return QVNT_OK;
}

}  // extern "C"
