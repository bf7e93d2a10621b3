// The device register (the reference's QReg, src/register/quant.rs:103-108)
// and the launch helpers shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "common.hpp"

#define QVNT_HIP(call)                                                                       \
    do {                                                                                     \
        hipError_t qvnt_e_ = (call);                                                         \
        if (qvnt_e_ != hipSuccess)                                                           \
            return ::qvnt::fail(QVNT_E_HIP, std::string(#call) + ": " + hipGetErrorString(qvnt_e_)); \
    } while (0)

struct qvnt_reg {
    uint32_t q_num = 0;      // qubits of the register
    uint32_t ns = 0;         // storage qubits: log2(len) = max(q_num, 3)
    uint64_t len = 0;        // stored amplitudes, max(2^q_num, 8) (quant.rs:15,116)
    qvnt_prec prec = QVNT_F64;
    int device = 0;
    void* psi = nullptr;     // device amplitudes, interleaved complex
    hipStream_t stream = nullptr;

    // measurement workspace: stored levels of the |psi|^2 pairwise tree
    double* tree = nullptr;
    size_t tree_cap = 0;     // doubles
    double* dscratch = nullptr;      // small device scratch (draws, totals)
    uint64_t* iscratch = nullptr;    // small device scratch: shots_cap outcomes + the total
    size_t shots_cap = 0;

    // fused-form program (phases, micro-ops, coefficients) on the device
    void* prog = nullptr;
    size_t prog_cap = 0;
    uint32_t dense_min = 0;  // fused form: phases of >= this many ops run as dense 16x16 (0: never)

    // per-kernel-class timing
    bool owns_memory = true;  // false: state, workspace and stream belong to a server slot
    bool timing = false;
    struct Pending { int kclass; hipEvent_t a, b; double bytes; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    qvnt_kstats stats[QVNT_KCLASS_COUNT] = {};

    size_t amp_bytes() const { return prec == QVNT_F64 ? 16 : 8; }
};

namespace qvnt {

// Brackets one launch with timing events when the register has timing on.
struct Timed {
    qvnt_reg* r;
    int kclass;
    double bytes;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(qvnt_reg* reg, int kc, double by);
    ~Timed();
};

// reg.hip: device blocks and streams that registers reuse.  Creating and
// destroying a stream and allocating and freeing the state and workspace
// cost more than the gates of a short-lived register (a 20-qubit register
// with a few gates and a measurement: most of ~290 us).  dev_alloc returns a
// cached block of exactly `bytes` of the device (current) or a new one;
// dev_release keeps a block of at most kCacheBlock bytes while the
// device's cache holds at most kCacheTotal, and frees the rest.  The caller
// of dev_release guarantees that no queued work still uses the block (its
// stream synchronized).  An allocation that fails frees every cached block
// of the device and tries once more.  stream_acquire / stream_release do
// the same for idle streams (default priority, non-blocking).
constexpr size_t kCacheBlock = size_t(64) << 20;
constexpr size_t kCacheTotal = size_t(256) << 20;
int dev_alloc(int device, void** p, size_t bytes);
void dev_release(int device, void* p, size_t bytes);
int stream_acquire(int device, hipStream_t* s);
void stream_release(int device, hipStream_t s);

// plain.hip
// Validates a gate against a buffer of `len` amplitudes (bit counts, control
// overlap, qubits beyond the buffer); *noop is set for gates that are the
// identity there (Id, a control bit beyond the buffer).
int check_gate(uint64_t len, const qvnt_gate& g, bool* noop);
int launch_gate(qvnt_reg* r, const qvnt_gate& g);
int launch_init(qvnt_reg* r, uint64_t state);

// measure.hip
int tree_build(qvnt_reg* r, uint64_t count, double* total_out);
int tree_sample(qvnt_reg* r, uint64_t count, const double* u, size_t shots, uint64_t* out, double* total_out);
// doubles of tree workspace the draws over `count` weights need
size_t tree_doubles(uint64_t count);
// enqueues the tree levels and one draw per u_dev[k] into out_dev (device
// pointers; r->tree must hold tree_doubles(count)); no synchronisation
int sample_launch(qvnt_reg* r, uint64_t count, const double* u_dev, size_t shots, uint64_t* out_dev, double* total_dev);
int collapse(qvnt_reg* r, uint64_t outcome, uint64_t mask, bool renorm);
int probabilities(qvnt_reg* r, double inv_total, double* dst, uint64_t n);
// every amplitude of the buffer times `scale` (FP64 multiply)
int rescale(qvnt_reg* r, double scale);
int compare(qvnt_reg* a, qvnt_reg* b, double* out);

// fused.hip
int apply_fused(qvnt_reg* r, const qvnt_gate* g, size_t n, uint32_t max_qubits, uint32_t* passes);
// the steps of a fused apply on 2^ns (>= 2^kMinFused) amplitudes (host
// only; returns the tile size): the tile the run-time model picks when
// max_qubits = 0 and dense_min = 0 from 28 qubits, else max_qubits (0: the
// default) clamped to the precision's LDS tile and the buffer
namespace fusion { struct Step; }
unsigned compile_fused(const qvnt_gate* g, size_t n, unsigned ns, bool fp64, uint32_t max_qubits, uint32_t dense_min,
                       std::vector<fusion::Step>& steps);
// what apply_fused would compile for a buffer of 2^ns amplitudes (host
// only): out = {passes, tile qubits, summed op cost, dense phases}
void fused_plan_stats(const qvnt_gate* g, size_t n, unsigned ns, bool fp64, uint32_t max_qubits, uint32_t dense_min,
                      double* out);
namespace fusion { struct Program; }
size_t program_bytes(const fusion::Program& prog);
void pack_program(const fusion::Program& prog, unsigned char* host);
int launch_program(qvnt_reg* r, const fusion::Program& prog, const void* dev);
namespace fusion { struct BItem; }
// one batched launch: `blocks` tiles of 2^t amplitudes (block -> item by map)
int launch_batch(hipStream_t stream, qvnt_prec prec, void* arena, const void* prog, const fusion::BItem* items,
                 const uint32_t* map, uint32_t t, uint32_t blocks);
namespace fusion { struct PUnit; }
// the persistent batched kernel: blocks resident at once for tiles up to
// 2^tmax, and its cooperative launch over `rounds` rounds of work units
// (units [round_begin[r], round_begin[r+1]) form round r; device pointers)
int batch_persist_grid(qvnt_prec prec, uint32_t tmax, uint32_t* blocks);
// (status: 1 + rounds device words, zero at the first launch: the kernel
// sets status[0] to 1 if its launch was not cooperative, and uses
// status[1 + r] as round r's claim counter, leaving it zero; graph
// non-null: capture the launch into *graph instead of running it)
int launch_batch_persist(hipStream_t stream, qvnt_prec prec, void* arena, const void* prog, const fusion::BItem* items,
                         const fusion::PUnit* units, const uint32_t* round_begin, uint32_t rounds, uint32_t tmax,
                         uint32_t blocks, uint32_t* status, hipGraph_t* graph);

}  // namespace qvnt
