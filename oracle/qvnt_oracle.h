/*
 * qvnt_oracle.h — CPU restatement of qvnt 0.4.4's state-vector algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity oracle and the CPU baseline;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker.  The product path (qvnt_amd) never links,
 * imports or falls back to it.
 *
 * Everything here restates the reference Rust code (paths relative to the
 * reference root, each function cites the file:line it follows):
 *   - operator builders: src/operator/mod.rs, single/{pauli,rotate,swap}.rs, multi/{mod,h,qft}.rs
 *   - the per-gate gather f(psi, idx): src/operator/atomic/<gate>.rs
 *   - the element loop and control predicate: atomic/dispatch.rs:32-67
 *   - MultiOp::apply ping-pong with the per-apply copy: multi/mod.rs:96-114
 *   - QReg::{with_state, apply, get_probabilities, get_absolute,
 *     collapse_mask, measure_mask}: src/register/quant.rs
 * It is compiled with -ffp-contract=off (Rust never fuses a*b+c).
 *
 * The measurement draw of the reference (rand 0.8.5 WeightedIndex on an
 * unseeded thread_rng, quant.rs:496-497) cannot be reproduced; the seeded
 * canonical-tree contract below replaces it (orc_tree_*).
 */
#ifndef QVNT_ORACLE_H
#define QVNT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/qvnt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_c { double re, im; } orc_c;

/* ---- operator builders (write into out[], return gate count or < 0) ---- */
long orc_op_id(qvnt_gate* out);
long orc_op_x(uint64_t a, qvnt_gate* out);
long orc_op_y(uint64_t a, qvnt_gate* out);
long orc_op_z(uint64_t a, qvnt_gate* out);
long orc_op_s(uint64_t a, qvnt_gate* out);
long orc_op_t(uint64_t a, qvnt_gate* out);
long orc_op_rx(double phase, uint64_t a, qvnt_gate* out);
long orc_op_ry(double phase, uint64_t a, qvnt_gate* out);
long orc_op_rz(double phase, uint64_t a, qvnt_gate* out);
long orc_op_rxx(double phase, uint64_t ab, qvnt_gate* out);
long orc_op_ryy(double phase, uint64_t ab, qvnt_gate* out);
long orc_op_rzz(double phase, uint64_t ab, qvnt_gate* out);
long orc_op_phi(double phase, uint64_t a, qvnt_gate* out);
long orc_op_swap(uint64_t ab, qvnt_gate* out);
long orc_op_sqrt_swap(uint64_t ab, qvnt_gate* out);
long orc_op_i_swap(uint64_t ab, qvnt_gate* out);
long orc_op_sqrt_i_swap(uint64_t ab, qvnt_gate* out);
long orc_op_h(uint64_t a, qvnt_gate* out);
long orc_op_u1(double lam, uint64_t a, qvnt_gate* out);
long orc_op_u2(double phi, double lam, uint64_t a, qvnt_gate* out);
long orc_op_u3(double the, double phi, double lam, uint64_t a, qvnt_gate* out);
long orc_op_qft(uint64_t a, qvnt_gate* out);
long orc_op_qft_swapped(uint64_t a, qvnt_gate* out);
long orc_op_matrix1(uint64_t a, const double m[8], qvnt_gate* out);
long orc_op_matrix2(uint64_t a, uint64_t b, const double m[32], qvnt_gate* out);
/* Applicable::c on a queue (multi/mod.rs:125-132); 0 or QVNT_E_CTRL_OVERLAP. */
int  orc_op_c(qvnt_gate* g, size_t n, uint64_t c);
/* Applicable::dgr on a queue (multi/mod.rs:120-123). */
void orc_op_dgr(qvnt_gate* g, size_t n);
uint64_t orc_op_act_on(const qvnt_gate* g, size_t n);

/* ---- register (quant.rs:103-108) ---- */
typedef struct orc_reg {
    orc_c*   psi;
    uint64_t len;      /* max(2^q_num, 8) */
    uint32_t q_num;
    uint64_t q_mask;
    int      threads;  /* 1 = threading::Single, >1 = Multi(n) (OpenMP) */
} orc_reg;

orc_reg* orc_reg_with_state(uint32_t q_num, uint64_t state);
void     orc_reg_free(orc_reg* r);
double*  orc_reg_psi(orc_reg* r);
uint64_t orc_reg_len(const orc_reg* r);
/* QReg::num_threads (quant.rs:186-200): 0 ok, -1 for the reference's None. */
int      orc_reg_num_threads(orc_reg* r, int threads);
/* QReg::apply with a MultiOp (quant.rs:376-395 -> multi/mod.rs:96-114). */
void     orc_reg_apply(orc_reg* r, const qvnt_gate* g, size_t n);
/* QReg::get_absolute / get_probabilities, reference summation order. */
double   orc_reg_get_absolute(const orc_reg* r);
void     orc_reg_get_probabilities(const orc_reg* r, double* out);
/* QReg::measure_mask under the seeded tree contract; *err = 0 or
 * QVNT_E_ZERO_NORM.  flags & QVNT_MEASURE_RENORM renormalises. */
uint64_t orc_reg_measure_mask(orc_reg* r, uint64_t mask, double u, uint32_t flags, int* err);

/* ---- measurement contract (replaces WeightedIndex + thread_rng) ----
 * w_i = re*re + im*im (no FMA).  node(0,j) = w_j;
 * node(l,j) = node(l-1,2j) + node(l-1,2j+1): the pairwise tree over
 * `count` (a power of two) weights.  The draw: x = u * node(top,0); walk
 * from the root; at children (L, R) go left iff x < L or R == 0, else
 * x -= L and go right; the leaf reached is r. */
double   orc_tree_total(const double* psi_interleaved, uint64_t count);
uint64_t orc_tree_sample(const double* psi_interleaved, uint64_t count, double u, int* err);

/* ---- reset, register algebra, readback (quant.rs:156-166, 202-229,
 * 245-371, 397-431, 513-594) ---- */
void     orc_reg_reset(orc_reg* r, uint64_t i_state);
void     orc_reg_reset_by_mask(orc_reg* r, uint64_t mask);
orc_reg* orc_reg_tensor(const orc_reg* a, const orc_reg* b);
orc_reg* orc_reg_combine(const orc_reg* a, const orc_reg* b);
orc_reg* orc_reg_combine_with_unitary(const orc_reg* a, const orc_reg* b, const double c[8]);
int      orc_reg_linear_composition(orc_reg* r, const orc_reg* other, const double c[4]);
int      orc_reg_set_num(orc_reg* r, uint32_t q_num);
void     orc_reg_get_polar(const orc_reg* r, double* out);
int      orc_reg_sample_all(const orc_reg* r, uint64_t shots, uint64_t seed, uint64_t stream, uint64_t* hist);

/* ---- Sym::finish (qasm/sym.rs:40-74) ---- */
enum { ORC_SEP_NOP = 0, ORC_SEP_MEASURE = 1, ORC_SEP_IF = 2, ORC_SEP_RESET = 3 };
typedef struct orc_step {
    const qvnt_gate* g;   /* the step's op (queue of SingleOps) */
    size_t n;
    uint32_t sep;         /* ORC_SEP_* */
    uint32_t pad;
    uint64_t a, b;        /* MEASURE: q_arg, c_arg; IF: c mask, value; RESET: q mask */
} orc_step;
int orc_run_program(orc_reg* r, const orc_step* steps, size_t n, const qvnt_gate* tail, size_t tail_n,
                    uint32_t measure_xor, uint64_t* cval, uint32_t c_num, const double* draws, size_t n_draws,
                    size_t* used);

/* SplitMix64-based seeded uniform (same definition as qvnt_uniform). */
uint64_t orc_splitmix64(uint64_t* state);
double   orc_uniform(uint64_t seed, uint64_t stream, uint64_t index);

#ifdef __cplusplus
}
#endif
#endif
