// Host-side compiler of the fused form (see fused.hip for the kernel).
//
//   1. lower every SingleOp to primitives: a 2x2 matrix on one qubit (M2),
//      a 4x4 matrix on two qubits (M4) or a diagonal phase table on one or
//      two qubits (D), each with the gate's control mask.  Multi-bit x/y/z/
//      s/t masks are tensor powers (atomic/x.rs etc.), so they split into
//      one primitive per bit;
//   2. multiply runs of uncontrolled one-qubit primitives on the same qubit
//      into one 2x2 (u3 = rz.ry.rz becomes one M2);
//   3. schedule passes greedily: a pass owns a set T of at most `tq` qubits
//      (always including qubits 0..2, so every tile row is whole 128-byte
//      lines in FP64; QVNT_ROW_BITS changes the count); walking the queue
//      in order, a primitive joins the pass if its
//      matrix qubits fit in T (diagonals and controls may sit anywhere) and
//      none of its qubits is held by an earlier primitive left out -- it
//      then commutes with everything it overtakes;
//   4. split each pass into phases: a phase holds four of the tile's qubits
//      in registers (16 amplitudes per thread); its matrix primitives act
//      on those, diagonals on any qubit; each diagonal then moves to the
//      phase within its reach where its qubits are registers (or uniform
//      over a wave);
//   5. (dense_min > 0, tiles of >= 2^10 amplitudes) a phase of at least
//      dense_min primitives that all act on its four register qubits only
//      (operands and controls) becomes one dense 16x16 unitary, applied as
//      a complex GEMM on the matrix cores;
//   6. other phases: 2x2s split into diagonal * real rotation * diagonal,
//      diagonals merged within their windows, runs of diagonal / rotations
//      / diagonal folded into one structured op, diagonals on bits outside
//      the tile reduced to a per-tile scalar -- kept only where the op-cost
//      estimate says the phase gets cheaper.
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace qvnt {
namespace fusion {

constexpr int kRegBits = 4;                // register qubits per phase
constexpr int kRegs = 1 << kRegBits;       // amplitudes per thread
constexpr int kMaxTile = 13;               // t: at most 2^13 amplitudes per tile (64 KiB of LDS in FP32)
// the largest tile whose amplitudes fit the kernel's 64 KiB LDS tile
constexpr unsigned max_tile(bool fp64) { return fp64 ? 12u : 13u; }
// the default tile cap: 2^12 amplitudes in both precisions (an FP32 tile of
// 2^13 is one 512-thread block per CU and measured slower on large states)
constexpr unsigned kDefaultTile = 12;
constexpr int kMinFused = 6;               // below 2^6 amplitudes the plain kernels run
constexpr int kMinDenseTile = 10;          // dense phases need whole 64-lane waves: 2^(t-4) >= 64
// From 2^20 amplitudes a pass is bandwidth-bound (16 MiB FP64), and a pass
// that would hold a single SingleOp runs as that SingleOp's plain kernel
// (same or fewer bytes, no LDS round trips); below, the single-register
// fused form keeps every pass fused, like the batched form.
constexpr unsigned kSinglePlainQubits = 20;

// ---------------------------------------------------------- the program
// M2: 2x2 on one register bit; M4: 4x4 on two; D: diagonal table;
// P4: a 4x4 that only pairs index r with r ^ 3 (rxx, ryy, the swap family):
// each output is two products instead of four; A2: an anti-diagonal 2x2
// (y, or x with a phase); X2: the pure swap (x, cx); T2: the real middle
// factor of a 2x2 split into diagonal * real * diagonal (see split_pairs
// in fusion.cpp), one real coefficient a: out0 = x0 - a x1,
// out1 = x1 + a x0, two FMAs per amplitude instead of the eight of a
// complex 2x2; S: one structured block, a register diagonal, T2 on each of
// the four register bits (a = 0, an exact identity, where a bit has none)
// and a second register diagonal, in one op without branches; G: a
// diagonal whose phase depends on bits outside the tile only (a scalar per
// tile): the kernel multiplies the phase's G factors into one uniform
// scalar and scales the amplitudes once, after the phase's other ops;
// PH: a phase on the 8 amplitudes whose register bit j is 1 (the |11>
// factor of a two-qubit diagonal, its other qubit a control); P4R: a P4
// whose pairing coefficients are real on the diagonal and imaginary off
// it (rxx, ryy), out_r = c_r x_r + i s_r x_(r^3): four real operations per
// output instead of eight.
enum : uint8_t { OP_M2 = 0, OP_M4 = 1, OP_D = 2, OP_P4 = 3, OP_A2 = 4, OP_X2 = 5, OP_T2 = 6, OP_S = 7, OP_G = 8,
                 OP_PH = 9, OP_P4R = 10 };

// All program fields are 32/64-bit and uniform across a block, so the
// kernel reads them with scalar loads into SGPRs.
struct FOp {                // 64 bytes
    uint32_t type;          // OP_*
    uint32_t sel;           // M2/A2/X2/T2/PH: register bit j0; M4/P4/P4R: 4*j0 + j1 (j0 < j1);
                            // S: bit 0 / bit 1 set when its first / second diagonal is applied
    uint32_t pc;            // matrix ops: register indices that satisfy the register controls (0xFFFF: all)
    uint32_t lctrl;         // controls on lane bits (local positions)
    uint32_t la, lb;        // D: operand a / b when it is a lane bit (local position mask)
    uint32_t coef;          // offset (in doubles) into the coefficient array
    uint32_t dmode;         // D: 0 per-lane variant table, 1 ratio table (see below)
    uint64_t gctrl;         // controls on bits outside the tile (global positions)
    uint64_t ga, gb;        // D: operand a / b when it lies outside the tile; T2, PH: the op's one
                            // coefficient inline, real part in ga and imaginary part in gb
                            // (double bit patterns), so the kernel needs no dependent load for it
    uint64_t pad2;
};
// D coefficients: 4 variants (k = [a is a non-register 1] | 2 [b is a
// non-register 1]) of the 16 per-register phases, register controls
// folded in (phase 1 where they fail).  dmode 1 (exactly one non-register
// operand, and its bit is the same for all 64 lanes of a wave: a bit
// outside the tile or a thread-index bit >= 6): the operand's variant slot
// (1 or 2) holds the ratio T_v / T_0, applied after T_0 by the waves whose
// operand bit is set -- every table load is uniform.  dmode 0: each lane
// reads its own variant's table.  M2: 4 complex; M4: 16 complex;
// P4: (u[r][r], u[r][r^3]) for r = 0..3; P4R: (c_r, s_r) as one complex
// per r; A2: (u01, u10); X2: none;
// T2: a (inline in the op); S: 16 complex (first diagonal),
// 4 complex whose real parts are the T2 coefficients of register bits
// 0..3, 16 complex (second diagonal); G: a D table (entry 0 of variant k is
// its factor); PH: one complex (inline in the op).

struct FPhase {
    uint32_t op_begin, op_end;
    uint32_t rpos[kRegBits];  // local positions of the register bits, ascending
    uint32_t sroff[kRegs];    // swizzled LDS offset of register i (the swizzle is XOR-linear)
    uint32_t dense;           // 1: the phase is one 16x16 unitary on its register bits (matrix cores)
    uint32_t ucoef;           // dense: offset (in doubles) of U, row-major complex
};

struct FPass {
    uint32_t t;               // tile qubits
    uint32_t nthreads;        // 2^(t - 4)
    uint32_t phase_begin, phase_end;
    uint32_t direct_first;    // 1: the first phase's registers are local bits t-4..t-1
    uint32_t direct_last;     // 1: same for the last phase (0 when permuted)
    uint32_t tpos[kMaxTile];  // global positions of the local bits, ascending
    uint64_t hoff[kRegs];     // global offset of local bits t-4..t-1 = j
    // permuted: the pass ends with a permutation of its qubits (a relabeled
    // swap put back): the thread that stores local index L = tid + j * 2^(t-4)
    // reads it from LDS offset (XOR over tid's set bits k of pimg[k]) ^ pout[j]
    uint32_t permuted;
    uint32_t pimg[kMaxTile - kRegBits];
    uint32_t pout[kRegs];
    // tpos again, one byte each (byte k of the 16 = tpos[k], 0 past t):
    // four scalar registers for the kernel's loops over the local bits
    uint32_t tpack[4];
};

// One register's pass in a batched launch: byte offsets into the program
// buffer and the register's amplitude offset in the batch arena.
struct BItem {              // 32 bytes
    uint64_t psi_off;       // amplitudes
    uint32_t pass_off;      // bytes: FPass
    uint32_t phase_off;     // bytes: the register program's FPhase array
    uint32_t op_off;        // bytes: its FOp array
    uint32_t coef_off;      // bytes: its coefficients
    uint32_t block0;        // first block of the item in its launch
    uint32_t tiles;         // tiles of the pass (k_batch: batch_tiles per block)
};

// Tiles one block of k_batch runs, one after the other.  A block's tiles
// share its register pass, so the program's headers and coefficients are
// in the scalar cache after the first; that pays in FP64 (two waves per
// SIMD, little latency hidden: 4096-register batch 48.9 ms per run with one
// tile per block, 47.2 with two or four) but not in FP32 (four waves per
// SIMD: 27.1 ms with one, 32.1 with two).
constexpr uint32_t batch_tiles(bool fp64) { return fp64 ? 2u : 1u; }

// LDS swizzle: bits 4..7 of the local index are XORed into bits 0..3, so
// 16 lanes that differ in bits 4..7 hit 16 different 16-byte bank groups.
inline uint32_t lds_swizzle(uint32_t e) { return e ^ ((e >> 4) & 15u); }

// A work unit of the persistent batched kernel: tiles [tile0, tile_end) of
// one register pass (items[item]), at most 2^(tmax - t) of them.
struct PUnit {              // 16 bytes
    uint32_t item;
    uint32_t tile0, tile_end;
    uint32_t pad;
};

struct Program {
    std::vector<FPass> passes;
    std::vector<FPhase> phases;
    std::vector<FOp> ops;
    std::vector<double> coef;
};

// One step of a fused apply: a compiled program, or one gate the fused
// kernel does not take (run by the plain kernels).
struct Step {
    bool plain = false;
    qvnt_gate gate{};
    Program prog;
};

// Compiles a validated queue for a buffer of 2^ns amplitudes with tiles of
// at most tq qubits.  Identity gates (Id, a control beyond the buffer) are
// dropped.  With single_plain, a pass that would hold one SingleOp and
// nothing else becomes a plain step for that SingleOp.  With parallel, the
// passes of a long queue compile on several host threads (the output is
// the same); callers that already compile many queues at once leave it off.
// amp_bytes (16: FP64, 8: FP32) sets how many low qubits every pass holds,
// so that each tile row is one 128-byte line.
void plan(const qvnt_gate* g, size_t n, unsigned ns, unsigned tq, std::vector<Step>& steps, unsigned dense_min = 0,
          bool single_plain = false, bool parallel = false, unsigned amp_bytes = 16);

// Sum over the program's ops of the compiler's estimate of vector
// instructions per thread (arithmetic per output amplitude plus a fixed
// cost per op).
uint64_t program_cost(const Program& prog);

// The queue compiled (without dense phases) for each tile size in
// kAutoTiles that fits the buffer, keeping the steps with the lowest
// estimated run time: passes x a_t + program_cost x b_t (a plain step
// counted as a pass), with a_t, b_t per tile size and precision fitted to
// measured fused applies on MI355X.  Larger tiles need fewer passes;
// smaller tiles run each op faster (more, smaller blocks per CU hide the
// barriers between phases).  Returns the tile size kept.  (13 only in
// FP32: an FP64 tile stops at 12, max_tile.)
constexpr unsigned kAutoTiles[] = {10, 11, 12, 13};
// From 2^28 amplitudes a pass takes milliseconds, more than compiling the
// candidates costs; single-register applies choose the tile from there.
constexpr unsigned kAutoTileQubits = 28;
unsigned plan_auto(const qvnt_gate* g, size_t n, unsigned ns, bool fp64, std::vector<Step>& steps,
                   bool single_plain = false, bool parallel = false);

}  // namespace fusion
}  // namespace qvnt
