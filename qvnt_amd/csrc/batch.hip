// Batched uneven registers (no reference counterpart: the reference runs
// one register at a time).
//
// Many independent registers of different sizes live in one HBM arena.
// Each register's circuit is compiled by the fused-form compiler into
// passes over tiles of at most `cap(n)` qubits; pass k of every register
// that has one forms round k, and a round is one k_batch launch per tile
// size present in it (blocks map to (register, tile) through a block map).
// Rounds run in order on the batch's stream; the whole sequence can be
// captured once in a hipGraph and replayed.  The tile cap per qubit count
// is picked by an autotuner that times every candidate cap on the batch's
// own registers of that size.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <new>
#include <thread>
#include <vector>

#include "fusion.hpp"
#include "reg.hpp"

using qvnt::fail;
using qvnt::fusion::BItem;
using qvnt::fusion::FPass;
using qvnt::fusion::FPhase;
using qvnt::fusion::FOp;
using qvnt::fusion::Program;
using qvnt::fusion::PUnit;
using qvnt::fusion::Step;

namespace {

constexpr unsigned kMinQubits = qvnt::fusion::kMinFused;
constexpr unsigned kMaxQubits = 30;
constexpr unsigned kTuneCaps[] = {10, 11, 12};  // FP64; FP32 tiles are one qubit larger (tune_cap)
constexpr unsigned kTuneDense[] = {0, 3, 4};

size_t a256(size_t x) { return (x + 255) & ~size_t(255); }

__global__ void k_zero(uint4* __restrict__ p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// amplitude state_k of register k := 1 (arena already zeroed)
template <class R>
__global__ void k_basis(R* __restrict__ arena, const uint64_t* __restrict__ off, const uint64_t* __restrict__ state,
                        uint32_t count) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) arena[2 * (off[k] + state[k])] = (R)1;
}

// One register's compiled circuit.
struct RegPlan {
    Program prog;
    size_t bytes = 0;   // packed size: passes, phases, ops, coefficients
};

// Packs a register program as [passes | phases | ops | coefficients] (each
// 256-byte aligned) into dst (null: size only); returns the packed size.
size_t pack(const RegPlan& rp, unsigned char* dst, size_t* pass_off, size_t* phase_off, size_t* op_off, size_t* coef_off) {
    const Program& p = rp.prog;
    size_t o = 0;
    *pass_off = o;
    o = a256(o + p.passes.size() * sizeof(FPass));
    *phase_off = o;
    o = a256(o + p.phases.size() * sizeof(FPhase));
    *op_off = o;
    o = a256(o + p.ops.size() * sizeof(FOp));
    *coef_off = o;
    o = a256(o + p.coef.size() * sizeof(double) + 64);
    if (dst) {
        std::memset(dst, 0, o);
        if (!p.passes.empty()) std::memcpy(dst + *pass_off, p.passes.data(), p.passes.size() * sizeof(FPass));
        if (!p.phases.empty()) std::memcpy(dst + *phase_off, p.phases.data(), p.phases.size() * sizeof(FPhase));
        if (!p.ops.empty()) std::memcpy(dst + *op_off, p.ops.data(), p.ops.size() * sizeof(FOp));
        if (!p.coef.empty()) std::memcpy(dst + *coef_off, p.coef.data(), p.coef.size() * sizeof(double));
    }
    return o;
}

size_t packed_bytes(const RegPlan& rp) {
    size_t a, b, c, d;
    return pack(rp, nullptr, &a, &b, &c, &d);
}

struct Launch {
    uint32_t t = 0;
    uint32_t blocks = 0;
    size_t items_off = 0;  // bytes into the launch buffer
    size_t map_off = 0;
};

// The launches of one compiled set of registers, and their device buffers.
struct Schedule {
    std::vector<Launch> launches;
    uint32_t rounds = 0;
    uint64_t passes = 0, tiles = 0, phases = 0, dense_phases = 0;
    double dense_amps = 0.0;  // amplitudes transformed by dense phases per run
    double bytes = 0.0;    // algorithmic HBM bytes of one run: 2 * amp * 2^n per pass
    void* dprog = nullptr;
    void* dlaunch = nullptr;
    // the persistent form: work units of every round (k_batch_persist)
    uint32_t tmax = 0;             // largest tile (the block holds 2^(tmax-4) threads)
    uint32_t persist_blocks = 0;   // resident blocks of the cooperative launch (0: not available)
    size_t p_items_off = 0, p_units_off = 0, p_rounds_off = 0, p_status_off = 0;   // bytes into dpersist
    void* dpersist = nullptr;
    void release() {
        if (dprog) (void)hipFree(dprog);
        if (dlaunch) (void)hipFree(dlaunch);
        if (dpersist) (void)hipFree(dpersist);
        dprog = dlaunch = dpersist = nullptr;
        persist_blocks = 0;
        launches.clear();
    }
};

}  // namespace

struct qvnt_batch {
    uint32_t count = 0;
    qvnt_prec prec = QVNT_F64;
    int device = 0;
    std::vector<uint32_t> q;
    std::vector<uint64_t> off;     // amplitude offset of each register
    std::vector<uint64_t> init;    // initial basis state of each register
    void* arena = nullptr;
    uint64_t arena_amps = 0;
    uint64_t* dmeta = nullptr;     // device copies of off, init
    hipStream_t stream = nullptr;
    Schedule sched;
    uint64_t single_ops = 0;
    uint32_t caps[kMaxQubits + 1] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // the persistent launch captured into a graph; persist_form: 0 not run
    // yet, 1 the graph, 2 direct cooperative launches (the capture failed or
    // its launch was not cooperative)
    hipGraph_t pgraph = nullptr;
    hipGraphExec_t pexec = nullptr;
    int persist_form = 0;
    qvnt_reg view;                 // non-owning register on one batch member (readback, sampling)

    size_t amp() const { return prec == QVNT_F64 ? 16 : 8; }
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (pexec) (void)hipGraphExecDestroy(pexec);
        if (pgraph) (void)hipGraphDestroy(pgraph);
        exec = pexec = nullptr;
        graph = pgraph = nullptr;
        persist_form = 0;
    }
};

namespace {

int reset(qvnt_batch* b) {
    const uint64_t units = b->arena_amps * b->amp() / 16;
    const uint64_t blocks = std::min<uint64_t>((units + 255) / 256, 1u << 16);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, b->stream, static_cast<uint4*>(b->arena), units);
    QVNT_HIP(hipGetLastError());
    const unsigned g = (b->count + 255) / 256;
    if (b->prec == QVNT_F64)
        hipLaunchKernelGGL(k_basis<double>, dim3(g), dim3(256), 0, b->stream, static_cast<double*>(b->arena), b->dmeta,
                           b->dmeta + b->count, b->count);
    else
        hipLaunchKernelGGL(k_basis<float>, dim3(g), dim3(256), 0, b->stream, static_cast<float*>(b->arena), b->dmeta,
                           b->dmeta + b->count, b->count);
    QVNT_HIP(hipGetLastError());
    return QVNT_OK;
}

// Compiles the circuits of registers `sel` with tile cap caps[q] and dense
// threshold dense[q] (parallel over host threads) for amplitudes of
// amp_bytes.  All gates were validated by the caller.
void compile_plans(const std::vector<std::vector<qvnt_gate>>& gates, const std::vector<uint32_t>& q,
                   const std::vector<uint32_t>& sel, const uint32_t* caps, const uint32_t* dense, unsigned amp_bytes,
                   std::vector<RegPlan>& plans, int* bad) {
    std::atomic<size_t> next{0};
    std::atomic<int> err{-1};
    auto work = [&] {
        for (;;) {
            const size_t i = next.fetch_add(1);
            if (i >= sel.size()) return;
            const uint32_t k = sel[i];
            const unsigned ns = q[k];
            const unsigned tq = std::min<unsigned>(caps[ns], ns);
            std::vector<Step> steps;
            qvnt::fusion::plan(gates[k].data(), gates[k].size(), ns, tq, steps, dense[ns], false, false, amp_bytes);
            RegPlan& rp = plans[k];
            rp.prog = Program();
            // an all-fusable queue compiles to one program; a gate the fused
            // kernel does not take splits it (and has no batched form)
            if (steps.size() > 1 || (steps.size() == 1 && steps[0].plain)) err.store((int)k);
            else if (steps.size() == 1) rp.prog = std::move(steps[0].prog);
            rp.bytes = packed_bytes(rp);
        }
    };
    unsigned nt = std::thread::hardware_concurrency();
    nt = std::max(1u, std::min(nt, 16u));
    nt = (unsigned)std::min<size_t>(nt, std::max<size_t>(1, sel.size() / 16));
    std::vector<std::thread> pool;
    for (unsigned i = 1; i < nt; ++i) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    *bad = err.load();
}

// Packs the plans of `sel` and builds their rounds.
int build_schedule(qvnt_batch* b, const std::vector<RegPlan>& plans, const std::vector<uint32_t>& sel, Schedule& S) {
    S.release();
    S.rounds = 0;
    S.passes = S.tiles = S.phases = S.dense_phases = 0;
    S.dense_amps = 0.0;
    S.bytes = 0.0;
    size_t total = 0;
    std::vector<size_t> base(b->count, 0);
    for (uint32_t k : sel) {
        base[k] = total;
        total += plans[k].bytes;
        S.rounds = std::max<uint32_t>(S.rounds, (uint32_t)plans[k].prog.passes.size());
    }
    if (total >= (1ull << 32)) return fail(QVNT_E_OOM, "batched programs exceed 4 GiB");
    std::vector<unsigned char> host(total + 256);
    std::vector<size_t> poff(b->count), phoff(b->count), opoff(b->count), cfoff(b->count);
    for (uint32_t k : sel) {
        size_t a, p, o, c;
        pack(plans[k], host.data() + base[k], &a, &p, &o, &c);
        poff[k] = base[k] + a;
        phoff[k] = base[k] + p;
        opoff[k] = base[k] + o;
        cfoff[k] = base[k] + c;
    }
    // rounds: pass r of every register, one launch per tile size; and the
    // same work as units of the persistent kernel (up to 2^(tmax - t)
    // tiles of one register pass per block)
    S.tmax = 0;
    for (uint32_t k : sel)
        for (const FPass& fp : plans[k].prog.passes) S.tmax = std::max(S.tmax, fp.t);
    std::vector<BItem> p_items;
    std::vector<PUnit> p_units;
    std::vector<uint32_t> p_rounds(1, 0);
    std::vector<unsigned char> lhost;
    for (uint32_t r = 0; r < S.rounds; ++r) {
        std::map<uint32_t, std::vector<uint32_t>> by_t;
        for (uint32_t k : sel)
            if (r < plans[k].prog.passes.size()) by_t[plans[k].prog.passes[r].t].push_back(k);
        for (auto& kv : by_t) {
            Launch L;
            L.t = kv.first;
            std::vector<BItem> items;
            std::vector<uint32_t> map;
            for (uint32_t k : kv.second) {
                const FPass& fp = plans[k].prog.passes[r];
                BItem it{};
                it.psi_off = b->off[k];
                it.pass_off = (uint32_t)(poff[k] + r * sizeof(FPass));
                it.phase_off = (uint32_t)phoff[k];
                it.op_off = (uint32_t)opoff[k];
                it.coef_off = (uint32_t)cfoff[k];
                it.block0 = (uint32_t)map.size();
                const uint64_t tiles = 1ull << (b->q[k] - fp.t);
                if (map.size() + tiles >= (1ull << 31)) return fail(QVNT_E_ARG, "too many tiles in one round");
                it.tiles = (uint32_t)tiles;
                const uint32_t per_block = qvnt::fusion::batch_tiles(b->prec == QVNT_F64);
                for (uint64_t j = 0; j < tiles; j += per_block) map.push_back((uint32_t)items.size());
                items.push_back(it);
                const uint64_t per_unit = 1ull << (S.tmax - fp.t);
                for (uint64_t j = 0; j < tiles; j += per_unit)
                    p_units.push_back(PUnit{(uint32_t)p_items.size(), (uint32_t)j, (uint32_t)std::min(tiles, j + per_unit), 0u});
                p_items.push_back(it);
                S.passes += 1;
                S.tiles += tiles;
                S.phases += fp.phase_end - fp.phase_begin;
                for (uint32_t ph = fp.phase_begin; ph < fp.phase_end; ++ph)
                    if (plans[k].prog.phases[ph].dense) {
                        S.dense_phases += 1;
                        S.dense_amps += (double)(1ull << b->q[k]);
                    }
                S.bytes += 2.0 * (double)b->amp() * (double)(1ull << b->q[k]);
            }
            L.blocks = (uint32_t)map.size();
            L.items_off = a256(lhost.size());
            L.map_off = a256(L.items_off + items.size() * sizeof(BItem));
            lhost.resize(a256(L.map_off + map.size() * sizeof(uint32_t)));
            std::memcpy(lhost.data() + L.items_off, items.data(), items.size() * sizeof(BItem));
            std::memcpy(lhost.data() + L.map_off, map.data(), map.size() * sizeof(uint32_t));
            S.launches.push_back(L);
        }
        p_rounds.push_back((uint32_t)p_units.size());
    }
    uint32_t widest = 0;
    for (uint32_t r = 0; r < S.rounds; ++r) widest = std::max(widest, p_rounds[r + 1] - p_rounds[r]);
    uint32_t resident = 0;
    if (S.tmax >= (uint32_t)qvnt::fusion::kMinFused && qvnt::batch_persist_grid(b->prec, S.tmax, &resident) == QVNT_OK) {
        S.persist_blocks = std::min(resident, widest);
        // a multiple of 8: every XCD gets as many blocks (k_batch_persist
        // hands each XCD its own range of a round's units)
        if (S.persist_blocks >= 8) S.persist_blocks &= ~7u;
    } else {
        qvnt::clear_error();   // the persistent form is optional: the rounds run without it
    }
    S.p_items_off = 0;
    S.p_units_off = a256(p_items.size() * sizeof(BItem));
    S.p_rounds_off = a256(S.p_units_off + p_units.size() * sizeof(PUnit));
    S.p_status_off = a256(S.p_rounds_off + p_rounds.size() * sizeof(uint32_t));
    // (the status word and the rounds' claim counters start at zero: phost
    // is zeroed)
    std::vector<unsigned char> phost(a256(S.p_status_off + (1 + S.rounds) * sizeof(uint32_t)));
    if (!p_items.empty()) std::memcpy(phost.data() + S.p_items_off, p_items.data(), p_items.size() * sizeof(BItem));
    if (!p_units.empty()) std::memcpy(phost.data() + S.p_units_off, p_units.data(), p_units.size() * sizeof(PUnit));
    std::memcpy(phost.data() + S.p_rounds_off, p_rounds.data(), p_rounds.size() * sizeof(uint32_t));
    if (hipMalloc(&S.dprog, host.size()) != hipSuccess || hipMalloc(&S.dlaunch, lhost.size() + 256) != hipSuccess ||
        hipMalloc(&S.dpersist, phost.size()) != hipSuccess) {
        (void)hipGetLastError();
        S.release();
        return fail(QVNT_E_OOM, "device memory for the batched programs");
    }
    QVNT_HIP(hipMemcpyAsync(S.dprog, host.data(), host.size(), hipMemcpyHostToDevice, b->stream));
    if (!lhost.empty()) QVNT_HIP(hipMemcpyAsync(S.dlaunch, lhost.data(), lhost.size(), hipMemcpyHostToDevice, b->stream));
    QVNT_HIP(hipMemcpyAsync(S.dpersist, phost.data(), phost.size(), hipMemcpyHostToDevice, b->stream));
    QVNT_HIP(hipStreamSynchronize(b->stream));  // the host staging buffers die here
    return QVNT_OK;
}

int enqueue(qvnt_batch* b, const Schedule& S) {
    const unsigned char* L = static_cast<const unsigned char*>(S.dlaunch);
    for (const Launch& l : S.launches) {
        if (int rc = qvnt::launch_batch(b->stream, b->prec, b->arena, S.dprog,
                                        reinterpret_cast<const BItem*>(L + l.items_off),
                                        reinterpret_cast<const uint32_t*>(L + l.map_off), l.t, l.blocks))
            return rc;
    }
    return QVNT_OK;
}

// The whole schedule as one cooperative launch of the persistent kernel
// (graph non-null: build that launch as a graph instead).
int enqueue_persist(qvnt_batch* b, const Schedule& S, hipGraph_t* graph = nullptr) {
    if (!S.persist_blocks) return fail(QVNT_E_HIP, "the persistent batched form is not available on this device");
    unsigned char* P = static_cast<unsigned char*>(S.dpersist);
    return qvnt::launch_batch_persist(b->stream, b->prec, b->arena, S.dprog, reinterpret_cast<const BItem*>(P + S.p_items_off),
                                      reinterpret_cast<const PUnit*>(P + S.p_units_off),
                                      reinterpret_cast<const uint32_t*>(P + S.p_rounds_off), S.rounds, S.tmax,
                                      S.persist_blocks, reinterpret_cast<uint32_t*>(P + S.p_status_off), graph);
}

// Mode 2 of qvnt_batch_run.  The first call captures the cooperative
// launch into a graph, runs it and reads the status word: a launch the
// runtime did not make cooperative leaves every register untouched and the
// status set, so the graph is dropped, the status cleared and the run made
// with a direct cooperative launch, as every later one is.
int run_persist(qvnt_batch* b) {
    const Schedule& S = b->sched;
    if (b->persist_form == 1) {
        QVNT_HIP(hipGraphLaunch(b->pexec, b->stream));
        return QVNT_OK;
    }
    if (b->persist_form == 0) {
        b->persist_form = 2;
        hipGraph_t g = nullptr;
        if (enqueue_persist(b, S, &g) == QVNT_OK && hipGraphInstantiate(&b->pexec, g, nullptr, nullptr, 0) == hipSuccess) {
            b->pgraph = g;
            uint32_t status = 0;
            unsigned char* P = static_cast<unsigned char*>(S.dpersist);
            QVNT_HIP(hipGraphLaunch(b->pexec, b->stream));
            QVNT_HIP(hipMemcpyAsync(&status, P + S.p_status_off, sizeof status, hipMemcpyDeviceToHost, b->stream));
            QVNT_HIP(hipStreamSynchronize(b->stream));
            if (status == 0) {
                b->persist_form = 1;
                return QVNT_OK;
            }
            status = 0;
            QVNT_HIP(hipMemcpyAsync(P + S.p_status_off, &status, sizeof status, hipMemcpyHostToDevice, b->stream));
            QVNT_HIP(hipStreamSynchronize(b->stream));
            (void)hipGraphExecDestroy(b->pexec);
            (void)hipGraphDestroy(b->pgraph);
            b->pexec = nullptr;
            b->pgraph = nullptr;
        } else {
            if (g) (void)hipGraphDestroy(g);
            b->pexec = nullptr;
            (void)hipGetLastError();
            qvnt::clear_error();
        }
    }
    return enqueue_persist(b, S);
}

int check_batch(const qvnt_batch* b) {
    if (!b) return fail(QVNT_E_ARG, "null batch");
    int dev = 0;
    QVNT_HIP(hipGetDevice(&dev));
    if (dev != b->device) QVNT_HIP(hipSetDevice(b->device));
    return QVNT_OK;
}

// Times `S` on the batch stream (one warm-up run, then `runs` runs), ms per run.
int time_schedule(qvnt_batch* b, const Schedule& S, int runs, double* ms) {
    hipEvent_t e0, e1;
    QVNT_HIP(hipEventCreate(&e0));
    QVNT_HIP(hipEventCreate(&e1));
    int rc = enqueue(b, S);
    if (rc == QVNT_OK && hipEventRecord(e0, b->stream) != hipSuccess) rc = fail(QVNT_E_HIP, "event record");
    for (int i = 0; rc == QVNT_OK && i < runs; ++i) rc = enqueue(b, S);
    if (rc == QVNT_OK && hipEventRecord(e1, b->stream) != hipSuccess) rc = fail(QVNT_E_HIP, "event record");
    if (rc == QVNT_OK && hipEventSynchronize(e1) != hipSuccess) rc = fail(QVNT_E_HIP, "event sync");
    float f = 0.0f;
    if (rc == QVNT_OK && hipEventElapsedTime(&f, e0, e1) != hipSuccess) rc = fail(QVNT_E_HIP, "elapsed time");
    *ms = (double)f / runs;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

}  // namespace

extern "C" {

int qvnt_batch_new(uint32_t count, const uint32_t* q_nums, const uint64_t* states, qvnt_prec prec, qvnt_batch** out) {
    qvnt::clear_error();
    if (!out || !q_nums || count == 0) return fail(QVNT_E_ARG, "need registers and an output pointer");
    if (prec != QVNT_F64 && prec != QVNT_F32) return fail(QVNT_E_ARG, "unknown precision");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(QVNT_E_NO_DEVICE, "no HIP device visible");
    hipDeviceProp_t prop;
    int dev = 0;
    QVNT_HIP(hipGetDevice(&dev));
    QVNT_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return fail(QVNT_E_NO_DEVICE, "library is built for gfx950");
    qvnt_batch* b = new (std::nothrow) qvnt_batch;
    if (!b) return fail(QVNT_E_OOM, "host allocation failed");
    b->count = count;
    b->prec = prec;
    b->device = dev;
    b->q.assign(q_nums, q_nums + count);
    b->init.assign(count, 0);
    b->off.resize(count);
    uint64_t amps = 0;
    for (uint32_t k = 0; k < count; ++k) {
        if (b->q[k] < kMinQubits || b->q[k] > kMaxQubits) {
            delete b;
            return fail(QVNT_E_ARG, "batched registers hold 6..30 qubits");
        }
        if (states) b->init[k] = states[k] & ((1ull << b->q[k]) - 1u);
        b->off[k] = amps;
        amps += 1ull << b->q[k];
    }
    b->arena_amps = amps;
    for (unsigned n = 0; n <= kMaxQubits; ++n) b->caps[n] = qvnt::fusion::kDefaultTile;
    if (hipMalloc(&b->arena, amps * b->amp()) != hipSuccess || hipMalloc(&b->dmeta, 2 * count * sizeof(uint64_t)) != hipSuccess ||
        hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        qvnt_batch_free(b);
        return fail(QVNT_E_OOM, "device memory for the batch");
    }
    std::vector<uint64_t> meta(b->off);
    meta.insert(meta.end(), b->init.begin(), b->init.end());
    if (hipMemcpy(b->dmeta, meta.data(), meta.size() * sizeof(uint64_t), hipMemcpyHostToDevice) != hipSuccess) {
        qvnt_batch_free(b);
        return fail(QVNT_E_HIP, "batch metadata upload");
    }
    qvnt_reg& v = b->view;
    v.prec = prec;
    v.device = dev;
    v.stream = b->stream;
    v.owns_memory = false;
    if (int rc = reset(b)) {
        qvnt_batch_free(b);
        return rc;
    }
    *out = b;
    return QVNT_OK;
}

int qvnt_batch_reset(qvnt_batch* b) {
    if (int rc = check_batch(b)) return rc;
    return reset(b);
}

uint32_t qvnt_batch_count(const qvnt_batch* b) { return b ? b->count : 0; }

int qvnt_batch_compile(qvnt_batch* b, const qvnt_op* const* ops, uint32_t max_qubits, uint32_t dense_min,
                       qvnt_batch_info* info) {
    if (int rc = check_batch(b)) return rc;
    if (!ops) return fail(QVNT_E_ARG, "null operator array");
    std::vector<std::vector<qvnt_gate>> gates(b->count);
    uint64_t single = 0;
    for (uint32_t k = 0; k < b->count; ++k) {
        if (!ops[k]) return fail(QVNT_E_ARG, "null operator for register " + std::to_string(k));
        const uint64_t len = 1ull << b->q[k];
        for (const qvnt_gate& g : ops[k]->g) {
            bool noop = false;
            if (int rc = qvnt::check_gate(len, g, &noop))
                return fail(rc, "register " + std::to_string(k) + ": " + qvnt_last_error());
            if (!noop) gates[k].push_back(g);
        }
        single += ops[k]->g.size();
    }
    b->drop_graph();
    b->sched.release();
    std::vector<RegPlan> plans(b->count);
    std::vector<uint32_t> all(b->count);
    for (uint32_t k = 0; k < b->count; ++k) all[k] = k;
    constexpr size_t kCaps = sizeof(kTuneCaps) / sizeof(kTuneCaps[0]);
    constexpr size_t kDense = sizeof(kTuneDense) / sizeof(kTuneDense[0]);
    uint32_t caps[kMaxQubits + 1], dense[kMaxQubits + 1];
    double tune_ms[kCaps][kDense] = {}, single_ms[kMaxQubits + 1] = {};
    int bad = -1;
    auto no_batched_form = [&] {
        return fail(QVNT_E_ARG, "register " + std::to_string(bad) + " holds a gate with no batched form");
    };
    if (max_qubits) {
        const unsigned c = std::max(kMinQubits, std::min<unsigned>(max_qubits, qvnt::fusion::max_tile(b->prec == QVNT_F64)));
        for (unsigned n = 0; n <= kMaxQubits; ++n) {
            caps[n] = c;
            dense[n] = dense_min;
        }
    } else {
        // autotune on the whole batch (every candidate compiled and timed as
        // the batch will run: all sizes together, round by round):
        //  1. one tile cap and one dense threshold for every size, over the grid;
        //  2. for each size that fits one tile, its own cap = n (one pass), kept
        //     if the whole batch gets faster.
        // The timed runs evolve the registers: every way out of this branch,
        // an error included, sets them back to their initial basis states.
        struct ResetOnExit {
            qvnt_batch* b;
            bool armed = true;
            ~ResetOnExit() {
                if (armed) (void)reset(b);
            }
        } reset_guard{b};
        auto time_all = [&](const uint32_t* tc, const uint32_t* td, const std::vector<uint32_t>& recompile,
                            double* ms) -> int {
            compile_plans(gates, b->q, recompile, tc, td, (unsigned)b->amp(), plans, &bad);
            if (bad >= 0) return no_batched_form();
            Schedule S;
            if (int rc = build_schedule(b, plans, all, S)) return rc;
            const int rc = time_schedule(b, S, 2, ms);
            S.release();
            return rc;
        };
        double best = 1e300;
        uint32_t tc[kMaxQubits + 1], td[kMaxQubits + 1];
        auto tune_cap = [&](size_t ci) { return kTuneCaps[ci] + (b->prec == QVNT_F64 ? 0u : 1u); };
        for (size_t ci = 0; ci < kCaps; ++ci)
            for (size_t di = 0; di < kDense; ++di) {
                const unsigned cap_ci = tune_cap(ci);
                if (cap_ci < (unsigned)qvnt::fusion::kMinDenseTile && kTuneDense[di]) continue;  // no dense tiles
                for (unsigned n = 0; n <= kMaxQubits; ++n) {
                    tc[n] = cap_ci;
                    td[n] = kTuneDense[di];
                }
                double ms = 0.0;
                if (int rc = time_all(tc, td, all, &ms)) return rc;
                tune_ms[ci][di] = ms;
                if (ms < best) {
                    best = ms;
                    std::memcpy(caps, tc, sizeof(tc));
                    std::memcpy(dense, td, sizeof(td));
                }
            }
        // plans must match caps/dense again before the per-size trials
        compile_plans(gates, b->q, all, caps, dense, (unsigned)b->amp(), plans, &bad);
        if (bad >= 0) return no_batched_form();
        std::map<uint32_t, std::vector<uint32_t>> by_q;
        for (uint32_t k = 0; k < b->count; ++k) by_q[b->q[k]].push_back(k);
        for (auto& kv : by_q) {
            const unsigned n = kv.first;
            if (n > qvnt::fusion::max_tile(b->prec == QVNT_F64) || n <= caps[n]) continue;
            std::memcpy(tc, caps, sizeof(tc));
            tc[n] = n;
            double ms = 0.0;
            if (int rc = time_all(tc, dense, kv.second, &ms)) return rc;
            single_ms[n] = ms;
            if (ms < best) {
                best = ms;
                caps[n] = n;
            } else {
                compile_plans(gates, b->q, kv.second, caps, dense, (unsigned)b->amp(), plans, &bad);  // back to the kept plans
            }
        }
        reset_guard.armed = false;
        if (int rc = reset(b)) return rc;
    }
    compile_plans(gates, b->q, all, caps, dense, (unsigned)b->amp(), plans, &bad);
    if (bad >= 0) return no_batched_form();
    if (int rc = build_schedule(b, plans, all, b->sched)) return rc;
    std::memcpy(b->caps, caps, sizeof(caps));
    b->single_ops = single;
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->rounds = b->sched.rounds;
        info->launches = (uint32_t)b->sched.launches.size();
        info->passes = b->sched.passes;
        info->tiles = b->sched.tiles;
        info->bytes = b->sched.bytes;
        info->single_ops = single;
        info->phases = b->sched.phases;
        info->dense_phases = b->sched.dense_phases;
        info->dense_flops = 128.0 * b->sched.dense_amps;
        info->persist_blocks = b->sched.persist_blocks;
        for (unsigned n = 0; n <= kMaxQubits && n < 32; ++n) {
            info->tile_cap[n] = caps[n];
            info->dense_min[n] = dense[n];
            info->single_ms[n] = single_ms[n];
        }
        for (size_t ci = 0; ci < kCaps; ++ci) {
            info->tune_caps[ci] = kTuneCaps[ci] + (b->prec == QVNT_F64 ? 0u : 1u);
            for (size_t di = 0; di < kDense; ++di) info->tune_ms[ci][di] = tune_ms[ci][di];
        }
    }
    return QVNT_OK;
}

int qvnt_batch_run(qvnt_batch* b, int use_graph) {
    if (int rc = check_batch(b)) return rc;
    if (b->sched.launches.empty()) return QVNT_OK;
    // one persistent launch, replayed from a graph where its captured
    // launch is cooperative (run_persist)
    if (use_graph == 2) return run_persist(b);
    if (!use_graph) return enqueue(b, b->sched);
    if (!b->exec) {
        QVNT_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue(b, b->sched);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(QVNT_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        b->graph = g;
        QVNT_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    QVNT_HIP(hipGraphLaunch(b->exec, b->stream));
    return QVNT_OK;
}

int qvnt_batch_persist_form(const qvnt_batch* b, int* form) {
    if (!b || !form) return fail(QVNT_E_ARG, "null argument");
    *form = b->persist_form;
    return QVNT_OK;
}

int qvnt_batch_time(qvnt_batch* b, uint32_t runs, int use_graph, double* ms) {
    if (int rc = check_batch(b)) return rc;
    if (!ms || runs == 0) return fail(QVNT_E_ARG, "need runs > 0 and an output");
    hipEvent_t e0, e1;
    QVNT_HIP(hipEventCreate(&e0));
    QVNT_HIP(hipEventCreate(&e1));
    int rc = hipEventRecord(e0, b->stream) == hipSuccess ? QVNT_OK : fail(QVNT_E_HIP, "event record");
    for (uint32_t i = 0; rc == QVNT_OK && i < runs; ++i) rc = qvnt_batch_run(b, use_graph);
    if (rc == QVNT_OK && hipEventRecord(e1, b->stream) != hipSuccess) rc = fail(QVNT_E_HIP, "event record");
    if (rc == QVNT_OK && hipEventSynchronize(e1) != hipSuccess) rc = fail(QVNT_E_HIP, "event sync");
    float f = 0.0f;
    if (rc == QVNT_OK && hipEventElapsedTime(&f, e0, e1) != hipSuccess) rc = fail(QVNT_E_HIP, "elapsed time");
    *ms = (double)f / runs;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int qvnt_batch_sync(qvnt_batch* b) {
    if (int rc = check_batch(b)) return rc;
    QVNT_HIP(hipStreamSynchronize(b->stream));
    return QVNT_OK;
}

namespace {
int point_view(qvnt_batch* b, uint32_t k) {
    if (k >= b->count) return fail(QVNT_E_ARG, "register index out of range");
    qvnt_reg& v = b->view;
    v.q_num = b->q[k];
    v.ns = b->q[k];
    v.len = 1ull << b->q[k];
    v.psi = static_cast<unsigned char*>(b->arena) + b->off[k] * b->amp();
    return QVNT_OK;
}
}  // namespace

int qvnt_batch_get_state(qvnt_batch* b, uint32_t k, void* dst, size_t bytes) {
    if (int rc = check_batch(b)) return rc;
    if (int rc = point_view(b, k)) return rc;
    const size_t need = b->view.len * b->amp();
    if (!dst || bytes < need) return fail(QVNT_E_ARG, "destination smaller than the register");
    QVNT_HIP(hipMemcpyAsync(dst, b->view.psi, need, hipMemcpyDeviceToHost, b->stream));
    QVNT_HIP(hipStreamSynchronize(b->stream));
    return QVNT_OK;
}

int qvnt_batch_sample_mask(qvnt_batch* b, uint32_t k, uint64_t mask, const double* u, size_t shots, uint64_t* out,
                           double* total) {
    if (int rc = check_batch(b)) return rc;
    if (int rc = point_view(b, k)) return rc;
    if (shots == 0) return QVNT_OK;
    if (!u || !out) return fail(QVNT_E_ARG, "null draws or outcomes");
    if (shots > (1u << 20)) return fail(QVNT_E_ARG, "at most 2^20 shots per call");
    for (size_t i = 0; i < shots; ++i)
        if (!(u[i] >= 0.0 && u[i] < 1.0)) return fail(QVNT_E_ARG, "draws must lie in [0, 1)");
    double t = 0.0;
    if (int rc = qvnt::tree_sample(&b->view, b->view.len, u, shots, out, &t)) return rc;
    if (!(t > 0.0) || !std::isfinite(t)) return fail(QVNT_E_ZERO_NORM, "measurement of a zero-norm or non-finite state");
    mask &= b->view.len - 1u;
    for (size_t i = 0; i < shots; ++i) out[i] &= mask;
    if (total) *total = t;
    return QVNT_OK;
}

int qvnt_batch_free(qvnt_batch* b) {
    if (!b) return QVNT_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->drop_graph();
    b->sched.release();
    if (b->view.tree) (void)hipFree(b->view.tree);
    if (b->view.dscratch) (void)hipFree(b->view.dscratch);
    if (b->view.iscratch) (void)hipFree(b->view.iscratch);
    if (b->arena) (void)hipFree(b->arena);
    if (b->dmeta) (void)hipFree(b->dmeta);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return QVNT_OK;
}

}  // extern "C"
