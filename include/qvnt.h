/*
 * qvnt.h — C ABI of the MI355X state-vector engine (drop-in backend for the
 * qvnt 0.4.4 register / operator API).
 *
 * The reference is a Rust crate; its public surface on this path is
 *   op::{id,x,y,z,s,t,rx,ry,rz,rxx,ryy,rzz,swap,sqrt_swap,i_swap,sqrt_i_swap,
 *        h,u1,u2,u3,qft,qft_swapped}          src/operator/mod.rs:100-522
 *   Applicable::{c, dgr, act_on}              src/operator/applicable.rs:5-47
 *   MultiOp (queue, `*` appends, Debug list)  src/operator/multi/mod.rs:63-191
 *   QReg::{new, with_state, apply, get_probabilities, get_absolute,
 *          measure_mask, measure}             src/register/quant.rs:113-507
 * Every entry point below names the reference item it replaces.  A Rust
 * `-sys` crate binds these symbols one-to-one.
 *
 * Conventions
 *   - Every function returning `int` returns QVNT_OK (0) or a negative
 *     QVNT_E_* code; qvnt_last_error() then holds a message (thread-local).
 *   - Masks are uint64_t: bit k of a mask is qubit k is bit k of the
 *     amplitude index (little-endian), as in the reference.
 *   - Amplitudes are interleaved complex {re, im}: double for QVNT_F64,
 *     float for QVNT_F32.  A register of n qubits stores max(2^n, 8)
 *     amplitudes (src/register/quant.rs:15,116).
 *   - A register handle owns its device memory and one HIP stream; it is not
 *     thread-safe.  Distinct handles may be driven from different threads.
 *   - Operator handles are host objects (gate metadata only, as in the
 *     reference, src/operator/mod.rs:7-9).
 */
#ifndef QVNT_H
#define QVNT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define QVNT_OK              0
#define QVNT_E_INVALID_MASK (-1) /* wrong bit count: reference panics via expect
                                    (src/operator/mod.rs:153,172,...)        */
#define QVNT_E_CTRL_OVERLAP (-2) /* .c() overlaps act|ctrl: reference returns
                                    None (src/operator/single/mod.rs:109-118) */
#define QVNT_E_OOM          (-3) /* device memory / budget exhausted          */
#define QVNT_E_HIP          (-4) /* HIP runtime error                         */
#define QVNT_E_ZERO_NORM    (-5) /* measure on an all-zero/NaN state: reference
                                    panics in WeightedIndex (quant.rs:497)    */
#define QVNT_E_ARG          (-6) /* bad argument (null, size, mask beyond the
                                    register: reference index panic)          */
#define QVNT_E_NO_DEVICE    (-7) /* no usable gfx950 device                   */

/* ------------------------------------------------------------ gate kinds
 * One kind per reference atomic op (enum AtomicOpDispatch,
 * src/operator/atomic/dispatch.rs:82-105), same order, plus PHI, which the
 * reference README lists (README.md:69) but does not implement. */
enum qvnt_kind {
    QVNT_ID = 0,       /* atomic/id.rs           */
    QVNT_X,            /* atomic/x.rs            */
    QVNT_RX,           /* atomic/rx.rs           */
    QVNT_RXX,          /* atomic/rxx.rs          */
    QVNT_Y,            /* atomic/y.rs            */
    QVNT_RY,           /* atomic/ry.rs           */
    QVNT_RYY,          /* atomic/ryy.rs          */
    QVNT_Z,            /* atomic/z.rs            */
    QVNT_S,            /* atomic/s.rs            */
    QVNT_T,            /* atomic/t.rs            */
    QVNT_RZ,           /* atomic/rz.rs           */
    QVNT_RZZ,          /* atomic/rzz.rs          */
    QVNT_U1,           /* atomic/u1.rs (2x2 matrix on one qubit)  */
    QVNT_U2,           /* atomic/u2.rs (4x4 matrix on two qubits) */
    QVNT_H1,           /* atomic/h1.rs           */
    QVNT_H2,           /* atomic/h2.rs           */
    QVNT_SWAP,         /* atomic/swap.rs         */
    QVNT_ISWAP,        /* atomic/i_swap.rs       */
    QVNT_SQRT_SWAP,    /* atomic/sqrt_swap.rs    */
    QVNT_SQRT_ISWAP,   /* atomic/sqrt_i_swap.rs  */
    QVNT_PHI,          /* diag(1, e^{i theta}) on one qubit (new) */
    QVNT_KIND_COUNT
};

/* One gate == one reference SingleOp {act, ctrl, func}
 * (src/operator/single/mod.rs:42-47).  Plain old data, 312 bytes. */
typedef struct qvnt_gate {
    uint32_t kind;    /* enum qvnt_kind                                       */
    uint32_t dagger;  /* S, T, ISWAP, SQRT_SWAP, SQRT_ISWAP: 1 = daggered     */
    uint64_t act;     /* SingleOp::act (= AtomicOp::acts_on)                  */
    uint64_t a;       /* H2 / U2: a_mask; other kinds: == act                 */
    uint64_t b;       /* H2 / U2: b_mask; other kinds: 0                      */
    uint64_t ctrl;    /* SingleOp::ctrl                                       */
    double   re, im;  /* rotations: phase = (cos th/2, sin th/2) as the
                         reference stores it; PHI: (cos th, sin th)           */
    double   m[32];   /* U1: 4 complex (row-major 2x2); U2: 16 complex (4x4)  */
} qvnt_gate;

/* Amplitude precision of a register. */
typedef enum qvnt_prec { QVNT_F64 = 0, QVNT_F32 = 1 } qvnt_prec;

/* ------------------------------------------------------- errors / misc */
const char* qvnt_last_error(void);
/* Library version string, e.g. "qvnt-mi355x 0.1.0 (gfx950)". */
const char* qvnt_version(void);

/* Seeded uniform draw in [0, 1) used by measurement (the reference draws from
 * an unseeded thread_rng, src/register/quant.rs:496-497, so a seeded contract
 * replaces it): SplitMix64 of (seed, stream, index), top 53 bits * 2^-53. */
double qvnt_uniform(uint64_t seed, uint64_t stream, uint64_t index);

/* ------------------------------------------------------------ operators
 * qvnt_op == reference MultiOp (a queue of SingleOps).  Constructors take the
 * reference argument order (phase first, then mask).  Rotation and swap
 * constructors return QVNT_E_INVALID_MASK where the reference panics. */
typedef struct qvnt_op qvnt_op;

int  qvnt_op_id(qvnt_op** out);                                   /* op::id           mod.rs:113 */
int  qvnt_op_x(uint64_t a_mask, qvnt_op** out);                   /* op::x            mod.rs:136 */
int  qvnt_op_y(uint64_t a_mask, qvnt_op** out);                   /* op::y            mod.rs:199 */
int  qvnt_op_z(uint64_t a_mask, qvnt_op** out);                   /* op::z            mod.rs:258 */
int  qvnt_op_s(uint64_t a_mask, qvnt_op** out);                   /* op::s            mod.rs:277 */
int  qvnt_op_t(uint64_t a_mask, qvnt_op** out);                   /* op::t            mod.rs:297 */
int  qvnt_op_rx(double phase, uint64_t a_mask, qvnt_op** out);    /* op::rx           mod.rs:151 */
int  qvnt_op_ry(double phase, uint64_t a_mask, qvnt_op** out);    /* op::ry           mod.rs:214 */
int  qvnt_op_rz(double phase, uint64_t a_mask, qvnt_op** out);    /* op::rz           mod.rs:321 */
int  qvnt_op_rxx(double phase, uint64_t ab_mask, qvnt_op** out);  /* op::rxx          mod.rs:170 */
int  qvnt_op_ryy(double phase, uint64_t ab_mask, qvnt_op** out);  /* op::ryy          mod.rs:233 */
int  qvnt_op_rzz(double phase, uint64_t ab_mask, qvnt_op** out);  /* op::rzz          mod.rs:340 */
int  qvnt_op_phi(double phase, uint64_t a_mask, qvnt_op** out);   /* README.md:69 "phi" (absent in code) */
int  qvnt_op_swap(uint64_t ab_mask, qvnt_op** out);               /* op::swap         mod.rs:365 */
int  qvnt_op_sqrt_swap(uint64_t ab_mask, qvnt_op** out);          /* op::sqrt_swap    mod.rs:387 */
int  qvnt_op_i_swap(uint64_t ab_mask, qvnt_op** out);             /* op::i_swap       mod.rs:412 */
int  qvnt_op_sqrt_i_swap(uint64_t ab_mask, qvnt_op** out);        /* op::sqrt_i_swap  mod.rs:434 */
int  qvnt_op_h(uint64_t a_mask, qvnt_op** out);                   /* op::h            mod.rs:464, multi/h.rs:14-45 */
int  qvnt_op_u1(double lam, uint64_t a_mask, qvnt_op** out);      /* op::u1           mod.rs:472-474 */
int  qvnt_op_u2(double phi, double lam, uint64_t a_mask, qvnt_op** out);             /* op::u2 mod.rs:480-482 */
int  qvnt_op_u3(double the, double phi, double lam, uint64_t a_mask, qvnt_op** out); /* op::u3 mod.rs:499-501 */
int  qvnt_op_qft(uint64_t a_mask, qvnt_op** out);                 /* op::qft          mod.rs:511-513, multi/qft.rs:4-33 */
int  qvnt_op_qft_swapped(uint64_t a_mask, qvnt_op** out);         /* op::qft_swapped  mod.rs:520, multi/qft.rs:35-52 */
/* General-matrix atomics (atomic/u1.rs, atomic/u2.rs; crate-internal in the
 * reference, single/pauli.rs:32-39).  QVNT_E_INVALID_MASK unless the mask bit
 * counts are right and the matrix is unitary to 1e-12. */
int  qvnt_op_matrix1(uint64_t a_mask, const double m[8], qvnt_op** out);
int  qvnt_op_matrix2(uint64_t a_mask, uint64_t b_mask, const double m[32], qvnt_op** out);
/* An operator from raw gates (copied; "Id" gates are dropped as in
 * From<SingleOp> for MultiOp, multi/mod.rs:135-146). */
int  qvnt_op_from_gates(const qvnt_gate* gates, size_t count, qvnt_op** out);
int  qvnt_op_clone(const qvnt_op* op, qvnt_op** out);             /* Clone */
void qvnt_op_free(qvnt_op* op);
/* lhs *= rhs: appends rhs's queue (MulAssign, multi/mod.rs:166-170). */
int  qvnt_op_mul(qvnt_op* lhs, const qvnt_op* rhs);
/* Applicable::c in place (multi/mod.rs:125-132): QVNT_E_CTRL_OVERLAP (the
 * reference's None) leaves the operator unchanged. */
int  qvnt_op_c(qvnt_op* op, uint64_t c_mask);
/* Applicable::dgr in place: reverse the queue, dgr each (multi/mod.rs:120-123). */
int  qvnt_op_dgr(qvnt_op* op);
uint64_t qvnt_op_act_on(const qvnt_op* op);                       /* multi/mod.rs:116-118 */
size_t   qvnt_op_len(const qvnt_op* op);                          /* VecDeque::len */
/* Borrowed view of the gate queue, valid until the op is changed or freed. */
const qvnt_gate* qvnt_op_gates(const qvnt_op* op);
/* Debug string of the MultiOp, e.g. "[H3, H12, C8_H3, C2_SWAP9]"
 * (multi/mod.rs:90-94, single/mod.rs:70-76).  Writes at most cap bytes
 * (NUL-terminated); *need gets the full length + 1. */
int  qvnt_op_name(const qvnt_op* op, char* buf, size_t cap, size_t* need);

/* ------------------------------------------------------------- registers */
typedef struct qvnt_reg qvnt_reg;

/* QReg::new (quant.rs:113-125) == with_state(q_num, 0). */
int qvnt_reg_new(uint32_t q_num, qvnt_prec prec, qvnt_reg** out);
/* QReg::with_state (quant.rs:129-150): psi[state & q_mask] = 1. */
int qvnt_reg_with_state(uint32_t q_num, uint64_t state, qvnt_prec prec, qvnt_reg** out);
/* Drop (quant.rs: the Vec is freed).  Waits for the register's queued work;
 * its stream and its device blocks of at most 64 MiB then go to a
 * per-device cache (at most 64 idle streams and 256 MiB of blocks), from
 * which the next register of the same sizes takes them: creating a stream
 * and allocating cost more than the gates of a short-lived register.  An
 * allocation that fails empties the device's block cache and tries once
 * more. */
int qvnt_reg_free(qvnt_reg* reg);
/* Frees every cached block and idle stream of every device.  No counterpart
 * in the reference (it keeps no cache): for callers that share the GPU's
 * memory with another allocator and want the cache's memory back. */
int qvnt_cache_trim(void);
uint32_t qvnt_reg_num(const qvnt_reg* reg);                       /* QReg::num */
/* Number of stored amplitudes, max(2^n, 8). */
uint64_t qvnt_reg_len(const qvnt_reg* reg);
qvnt_prec qvnt_reg_prec(const qvnt_reg* reg);

/* QReg::apply(&MultiOp) (quant.rs:376-395): the plain form — one in-place
 * kernel per SingleOp, in queue order, on the register's stream.  Results
 * equal the reference's out-of-place gather (dispatch.rs:32-47). */
int qvnt_reg_apply(qvnt_reg* reg, const qvnt_op* op);
/* Same for a raw gate array (borrowed for the call). */
int qvnt_reg_apply_gates(qvnt_reg* reg, const qvnt_gate* gates, size_t count);

/* The fused form of QReg::apply (no reference counterpart: the reference
 * runs one gather pass per SingleOp).  The queue is scheduled into passes;
 * each pass reads and writes the state once and applies every gate it
 * absorbed while tiles of 2^t amplitudes (t = max_qubits, at most 12 for
 * FP64 and 13 for FP32) sit in LDS and registers.  max_qubits = 0: t = 12,
 * except from 28 qubits without dense phases, where the queue is compiled
 * for t = 10, 11 and 12 (and 13 in FP32) and the program with the lowest
 * estimated run time is kept (a model of passes and arithmetic fitted to measured runs; long
 * compute-heavy queues run faster in smaller tiles).  Gates are reordered only past gates on
 * disjoint qubits.  Same result as qvnt_reg_apply to rounding (FMA and
 * merged 2x2 products: about 1e-15 per gate).  *passes (may be NULL) gets
 * the number of passes over the state. */
int qvnt_reg_apply_fused(qvnt_reg* reg, const qvnt_op* op, uint32_t max_qubits, uint32_t* passes);
/* What qvnt_reg_apply_fused would compile for a register of q_num qubits of
 * precision prec with this max_qubits and dense-phase threshold (no
 * reference counterpart; host only, no device needed): out[0] = passes over
 * the state (plain launches counted), out[1] = tile qubits used (0: plain
 * kernels only), out[2] = the compiler's summed cost estimate of the ops
 * (vector instructions per thread), out[3] = dense phases. */
int qvnt_fused_plan_stats(const qvnt_op* op, uint32_t q_num, qvnt_prec prec, uint32_t max_qubits, uint32_t dense_min,
                          double* out);

/* QReg::get_absolute (quant.rs:458-466): sum of |psi_i|^2 over the buffer. */
int qvnt_reg_get_absolute(qvnt_reg* reg, double* out);
/* QReg::get_probabilities (quant.rs:434-454): p_i = |psi_i|^2 / total for
 * i < 2^n.  dst must hold 2^n doubles. */
int qvnt_reg_get_probabilities(qvnt_reg* reg, double* dst, uint64_t len);

/* QReg::measure_mask (quant.rs:490-501).  u is the caller's seeded uniform
 * draw in [0,1).  *out = r & mask where r is drawn with probability
 * |psi_r|^2 / total, then every amplitude with (idx ^ r) & mask != 0 is
 * zeroed.  flags & QVNT_MEASURE_RENORM rescales the kept amplitudes to unit
 * norm in FP64 (the reference leaves norm^2 = P(outcome)). */
#define QVNT_MEASURE_RENORM 1u
int qvnt_reg_measure_mask(qvnt_reg* reg, uint64_t mask, double u, uint32_t flags, uint64_t* out);
/* QReg::measure (quant.rs:505-507) == measure_mask(q_mask). */
int qvnt_reg_measure(qvnt_reg* reg, double u, uint32_t flags, uint64_t* out);
/* Non-destructive shots: out[k] = r_k & mask for draws u[k], state untouched
 * (each equals measure_mask on a copy of the register). */
int qvnt_reg_sample_mask(qvnt_reg* reg, uint64_t mask, const double* u, size_t shots, uint64_t* out);
/* qvnt_reg_sample_mask without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same outcomes as qvnt_reg_sample_mask for the same draws.  The measurement
 * tree, draws and outcomes live in one fixed 16 MiB device workspace per GPU,
 * held until the process ends: registers up to 30 qubits (fewer shots at 30);
 * a call that needs more returns QVNT_E_ARG.  All calls on one GPU share that
 * workspace: calls that overlap on one GPU (other threads, other registers)
 * race.  Synchronises the register's stream. */
int qvnt_reg_sample_mask_noalloc(qvnt_reg* reg, uint64_t mask, const double* u, size_t shots, uint64_t* out);

/* Device-side comparison of two registers of equal size (any precisions;
 * no reference counterpart): out = { Re<a|b>, Im<a|b>, |a|^2, |b|^2,
 * max_i |a_i - b_i| }, accumulated in FP64.  Infidelity is
 * 1 - |<a|b>|^2 / (|a|^2 |b|^2).  Synchronises both registers. */
int qvnt_reg_compare(qvnt_reg* a, qvnt_reg* b, double out[5]);

/* Copy-out / copy-in of the interleaved amplitudes (bytes = len * 2 *
 * sizeof(double or float)).  Blocking. */
int qvnt_reg_get_state(qvnt_reg* reg, void* dst, size_t bytes);
int qvnt_reg_set_state(qvnt_reg* reg, const void* src, size_t bytes);
/* Debug of QReg (quant.rs:603-623): "QReg { 0: Complex { re: 0.25, im: 0.0 },
 * ..., 7: ..., .. }" -- the first min(2^n, 8) amplitudes with Rust's f64
 * Debug formatting, ".." when the register holds more (FP32 amplitudes are
 * printed widened to f64).  Writes at most cap bytes (NUL-terminated);
 * *need gets the full length + 1.  Synchronises. */
int qvnt_reg_debug(qvnt_reg* reg, char* buf, size_t cap, size_t* need);
/* Wait for all work queued on the register's stream. */
int qvnt_reg_sync(qvnt_reg* reg);

/* ----------------------------------------------------------- serving
 * A continuous-batching request server (no reference counterpart).  A
 * request is one circuit on a fresh register of q_num qubits (basis state
 * init_state) followed by `shots` non-destructive draws of measure_mask on
 * the final state (each equals measure_mask on a copy of the register).
 * Requests are admitted as stream slots and space in the server's HBM arena
 * free up; requests in different slots overlap on the device; results do
 * not depend on what runs beside them. */
typedef struct qvnt_server qvnt_server;
typedef struct qvnt_request {
    uint64_t id;
    uint32_t q_num;
    uint32_t prec;            /* qvnt_prec                                     */
    uint64_t init_state;
    const qvnt_op* op;        /* copied at submit                              */
    uint32_t fused;           /* 1: fused form (qvnt_reg_apply_fused); 0: plain */
    uint32_t max_qubits;      /* fused tile qubits, 0 = default                */
    uint32_t dense_min;       /* fused: dense-phase threshold, 0 = none (qvnt_reg_set_dense) */
    uint32_t reserved;        /* 0                                             */
    uint64_t mask;            /* measured qubits                               */
    uint64_t shots;           /* 1 .. 2^20                                     */
    const double* draws;      /* shots uniform draws in [0,1), copied          */
    uint64_t* outcomes;       /* caller-owned; filled (r_k & mask) before the
                                 result is returned by qvnt_server_poll       */
} qvnt_request;
typedef struct qvnt_result {
    uint64_t id;
    int32_t status;           /* QVNT_OK, or QVNT_E_OOM if the request exceeds
                                 the arena, or another error code             */
    uint32_t passes;          /* kernel passes over the state                  */
    double total;             /* |psi|^2 of the final state                    */
    double t_submit, t_admit, t_done;  /* seconds on the server clock         */
} qvnt_result;
/* hbm_bytes: the arena (admission budget); slots: concurrent requests, one
 * HIP stream each.  More slots than the runtime's hardware queues (4 by
 * default) put several streams on one queue, where a small request's
 * kernels wait behind a large one's: 4 slots measured best. */
int qvnt_server_new(uint64_t hbm_bytes, uint32_t slots, qvnt_server** out);
/* QVNT_E_ARG if req->id belongs to a request submitted earlier and not yet
 * returned by qvnt_server_poll (its outcome buffer is still being filled). */
int qvnt_server_submit(qvnt_server* s, const qvnt_request* req);
/* Up to cap finished results; waits up to wait_ms for the first one. */
int qvnt_server_poll(qvnt_server* s, qvnt_result* out, uint32_t cap, uint32_t* n, int wait_ms);
int qvnt_server_pending(qvnt_server* s, uint32_t* queued, uint32_t* running, uint64_t* peak_bytes);
/* Seconds since the server started (the clock of t_submit/t_admit/t_done). */
double qvnt_server_now(qvnt_server* s);
/* Finishes in-flight requests, drops queued ones, frees everything. */
int qvnt_server_free(qvnt_server* s);

/* --------------------------------------------------------- instrumentation
 * Per-kernel-class device timing with HIP events recorded on the register's
 * stream around every launch (off by default). */
enum qvnt_kclass {
    QVNT_K_DIAG = 0,   /* phase-only diagonal, single pass                */
    QVNT_K_PAIR,       /* 2-amplitude orbit (pair mix / permutation)      */
    QVNT_K_QUAD,       /* 4-amplitude orbit (H2, U2)                      */
    QVNT_K_FUSED,      /* fused multi-gate tile pass                      */
    QVNT_K_REDUCE,     /* |psi|^2 tree reduction                          */
    QVNT_K_COLLAPSE,   /* measurement collapse / rescale                  */
    QVNT_K_OTHER,
    QVNT_KCLASS_COUNT
};
typedef struct qvnt_kstats {
    uint64_t launches;
    double   ms;       /* summed event time                               */
    double   bytes;    /* summed algorithmic bytes 2*s*2^(n-popcount(ctrl)) */
} qvnt_kstats;
int qvnt_reg_set_timing(qvnt_reg* reg, int enable);
/* Fused form (qvnt_reg_apply_fused): a phase of at least min_ops primitives
 * on four register qubits only, in tiles of >= 2^10 amplitudes, is folded
 * into one dense 16x16 unitary applied on the matrix cores (0: never, the
 * default).  No counterpart in the reference. */
int qvnt_reg_set_dense(qvnt_reg* reg, uint32_t min_ops);
/* Synchronises, folds pending events into the totals, copies
 * QVNT_KCLASS_COUNT entries into out, and (if reset) clears them. */
int qvnt_reg_kernel_stats(qvnt_reg* reg, qvnt_kstats* out, int reset);

/* ------------------------------------------------------------------ */
/* Reset, register algebra and readback (src/register/quant.rs). */
/* QReg::reset (quant.rs:202-205): zeros, psi[i_state & q_mask] = 1. */
int qvnt_reg_reset(qvnt_reg* reg, uint64_t i_state);
/* QReg::reset_by_mask (quant.rs:207-229): the whole register -> reset(0);
 * otherwise every amplitude with idx & mask != 0 is zeroed, then
 * QReg::normalize (quant.rs:397-414): norm = sqrt(sum |psi|^2); reset(0) if
 * norm <= 1e-15, unchanged if 1 - norm <= 1e-9, else scaled by 1/norm. */
int qvnt_reg_reset_by_mask(qvnt_reg* reg, uint64_t mask);
/* Mul for QReg (tensor_prod, quant.rs:330-371, 625-636): a new register of
 * a.q_num + b.q_num qubits, a on the low qubits.  a and b are kept. */
int qvnt_reg_tensor(qvnt_reg* a, qvnt_reg* b, qvnt_reg** out);
/* QReg::combine (quant.rs:245-268): equal sizes (>= 3 qubits; below, the
 * reference panics as its 8-element buffers do not fit as halves); a fills
 * the lower half of a register one qubit larger, b the upper half. */
int qvnt_reg_combine(qvnt_reg* a, qvnt_reg* b, qvnt_reg** out);
/* QReg::combine_with_unitary (quant.rs:271-308): m = [c00, c01, c10, c11]
 * (interleaved re, im), unitary by the reference's rule. */
int qvnt_reg_combine_with_unitary(qvnt_reg* a, qvnt_reg* b, const double m[8], qvnt_reg** out);
/* QReg::linear_composition (quant.rs:311-328): reg = reg * c0 + other * c1,
 * c = [c0.re, c0.im, c1.re, c1.im]; equal buffer lengths. */
int qvnt_reg_linear_composition(qvnt_reg* reg, qvnt_reg* other, const double c[4]);
/* QReg::set_num (quant.rs:156-166): the state is zero-extended or
 * truncated to 2^q_num amplitudes; nothing is renormalised (the reference's
 * reset branch is dead code). */
int qvnt_reg_set_num(qvnt_reg* reg, uint32_t q_num);
/* QReg::get_polar (quant.rs:417-431): (|psi_i|, arg psi_i) pairs, FP64,
 * for i < 2^n; dst holds 2 * 2^n doubles. */
int qvnt_reg_get_polar(qvnt_reg* reg, double* dst, uint64_t len);
/* QReg::sample_all (quant.rs:513-594) as an exact multinomial: `shots`
 * draws qvnt_uniform(seed, stream, k), k = 0..shots-1, under the measurement
 * contract, without collapse; hist[i] = count of outcome i, one entry per
 * buffer element (len >= max(2^n, 8)).  Replaces the reference's normal
 * approximation on an unseeded generator. */
int qvnt_reg_sample_all(qvnt_reg* reg, uint64_t shots, uint64_t seed, uint64_t stream, uint64_t* hist, uint64_t len);

/* ------------------------------------------------------------------ */
/* The QASM driver loop: Sym::finish (src/qasm/sym.rs:40-74) over the
 * (op, separator) list the interpreter builds (qasm/int/ext_op.rs:5-16). */
typedef enum {
    QVNT_SEP_NOP = 0,      /* apply op */
    QVNT_SEP_MEASURE = 1,  /* apply op; measure_mask(a) -> classical bits b (Sep::Measure(q_arg, c_arg)) */
    QVNT_SEP_IF = 2,       /* apply op only if the classical bits a read b (Sep::IfBranch(c, v)) */
    QVNT_SEP_RESET = 3     /* apply op; reset_by_mask(a) (Sep::Reset(q)) */
} qvnt_sep;
typedef struct {
    const qvnt_op* op;     /* may be null: no gates */
    uint32_t sep;          /* qvnt_sep */
    uint32_t pad;
    uint64_t a, b;
} qvnt_step;
/* Runs the steps, then `tail` (may be null), on reg and the classical
 * value *c_value of c_num bits.  Measurement: the i-th set bit of the
 * classical mask is set from (or, measure_xor, xored with) the i-th set bit
 * of the measured mask (MeasureOp::Set / Xor); each measurement with a
 * non-empty mask takes the next draw; *used = draws taken.  If: the
 * classical bits at the set bits of a, packed from bit 0 (CReg::get_by_mask,
 * class.rs:106-117), must equal b. */
int qvnt_reg_run_program(qvnt_reg* reg, const qvnt_step* steps, size_t n, const qvnt_op* tail, uint32_t measure_xor,
                         uint64_t* c_value, uint32_t c_num, const double* draws, size_t n_draws, size_t* used);

/* ------------------------------------------------------------------ */
/* Batched uneven registers (no counterpart in the reference, which runs
 * one register at a time): many independent registers of 6..30 qubits in
 * one device arena, each with its own circuit, compiled into fused passes
 * and run round by round (pass k of every register in one launch per tile
 * size), optionally as one captured hipGraph. */
typedef struct qvnt_batch qvnt_batch;
typedef struct {
    uint32_t rounds;          /* sequential rounds: the most passes of any register */
    uint32_t launches;        /* kernel launches per run */
    uint64_t passes;          /* register passes per run */
    uint64_t tiles;           /* tiles per run (k_batch runs two per workgroup in FP64) */
    uint64_t single_ops;      /* reference SingleOps of all circuits */
    double bytes;             /* algorithmic HBM bytes per run: 2 * amp * 2^n per pass */
    uint64_t phases;          /* phases per run (all passes) */
    uint64_t dense_phases;    /* of which dense 16x16 phases on the matrix cores */
    double dense_flops;       /* real matrix-core flops per run: 128 per amplitude per dense phase
                                 (a complex 16x16 as a real 32x32 GEMM) */
    uint32_t tile_cap[32];    /* tile cap used for registers of each qubit count */
    uint32_t dense_min[32];   /* dense-phase threshold used for each qubit count (0: none) */
    uint32_t tune_caps[3];    /* the tile caps the autotuner tried: 10, 11, 12 (FP64) or 11, 12, 13 (FP32) */
    uint32_t persist_blocks;  /* blocks of the persistent launch (qvnt_batch_run mode 2); 0: not available */
    double tune_ms[3][3];     /* autotuner: ms per run of the whole batch with one tile cap
                                 (tune_caps) and one dense threshold (0, 3, 4) for every size
                                 (0: not timed) */
    double single_ms[32];     /* autotuner: ms per run with registers of this size in one tile
                                 each, the others as chosen (0: not timed) */
} qvnt_batch_info;
/* Registers k = 0..count-1 of q_nums[k] qubits, each set to basis state
 * states[k] (null: all 0). */
int qvnt_batch_new(uint32_t count, const uint32_t* q_nums, const uint64_t* states, qvnt_prec prec, qvnt_batch** out);
uint32_t qvnt_batch_count(const qvnt_batch* b);
/* Every register back to its initial basis state (stream-ordered). */
int qvnt_batch_reset(qvnt_batch* b);
/* Compiles ops[k] (queue order, as qvnt_reg_apply) for register k.
 * max_qubits > 0 caps every tile at that many qubits (6..12 FP64, 6..13 FP32) and folds
 * phases of >= dense_min ops into dense 16x16 unitaries (0: none; see
 * qvnt_reg_set_dense).  max_qubits = 0 runs the autotuner instead: it times
 * the whole batch for every (tile cap, dense threshold) pair, then one tile
 * per register for each size that fits one, keeps the fastest, and leaves
 * the registers reset to their initial states. */
int qvnt_batch_compile(qvnt_batch* b, const qvnt_op* const* ops, uint32_t max_qubits, uint32_t dense_min,
                       qvnt_batch_info* info);
/* Queues one application of every compiled circuit.  graph = 0: the rounds
 * as separate launches (one per round and tile size); 1: the same captured
 * once (at the first such call after a compile) into a hipGraph and
 * replayed; 2: one persistent cooperative launch whose resident blocks walk
 * all rounds with a grid barrier between them, captured once into a
 * hipGraph and replayed (the first such call checks that the captured
 * launch ran cooperatively; if not, it and every later call launch the
 * kernel directly with hipLaunchCooperativeKernel, see
 * qvnt_batch_persist_form); fails with QVNT_E_HIP where
 * info->persist_blocks was 0.  All modes give the same amplitudes bit for
 * bit. */
int qvnt_batch_run(qvnt_batch* b, int graph);
/* How mode 2 of qvnt_batch_run runs (no reference counterpart): *form = 0
 * not run since the last compile, 1 the captured hipGraph, 2 direct
 * cooperative launches. */
int qvnt_batch_persist_form(const qvnt_batch* b, int* form);
/* Runs `runs` times between two device events (graph as qvnt_batch_run);
 * *ms = milliseconds per run. */
int qvnt_batch_time(qvnt_batch* b, uint32_t runs, int graph, double* ms);
int qvnt_batch_sync(qvnt_batch* b);
/* Register k's amplitudes (2^q interleaved complex of the batch precision). */
int qvnt_batch_get_state(qvnt_batch* b, uint32_t k, void* dst, size_t bytes);
/* Shots of measure_mask on register k without collapse (as
 * qvnt_reg_sample_mask); *total (may be null) = sum |psi|^2. */
int qvnt_batch_sample_mask(qvnt_batch* b, uint32_t k, uint64_t mask, const double* u, size_t shots, uint64_t* out,
                           double* total);
int qvnt_batch_free(qvnt_batch* b);

#ifdef __cplusplus
}
#endif
#endif /* QVNT_H */
