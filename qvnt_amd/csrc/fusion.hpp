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
//      lines); walking the queue in order, a primitive joins the pass if its
//      matrix qubits fit in T (diagonals and controls may sit anywhere) and
//      none of its qubits is held by an earlier primitive left out -- it
//      then commutes with everything it overtakes;
//   4. split each pass into phases: a phase holds four of the tile's qubits
//      in registers (16 amplitudes per thread); its matrix primitives act
//      on those, diagonals on any qubit.
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace qvnt {
namespace fusion {

constexpr int kRegBits = 4;                // register qubits per phase
constexpr int kRegs = 1 << kRegBits;       // amplitudes per thread
constexpr int kMaxTile = 12;               // t: at most 2^12 amplitudes per tile
constexpr int kMinFused = 6;               // below 2^6 amplitudes the plain kernels run

// ---------------------------------------------------------- the program
enum : uint8_t { OP_M2 = 0, OP_M4 = 1, OP_D = 2 };

struct FOp {
    uint8_t type;
    uint8_t j0, j1;        // register bits of the matrix qubits (M4: j0 < j1)
    uint8_t a_is_reg, b_is_reg;
    uint8_t pad[1];
    uint16_t pc;           // register-index pattern satisfying the register controls
    uint16_t pa, pb;       // D: bit patterns of operands a / b over register indices
    uint16_t lctrl;        // controls on lane (non-register tile) bits, local positions
    uint16_t la, lb;       // D: operand a / b as lane bits (local positions)
    uint16_t pad2;
    uint64_t gctrl;        // controls on bits outside the tile (global positions)
    uint64_t ga, gb;       // D: operands outside the tile (global positions)
    uint32_t coef;         // offset (in doubles) into the coefficient array
    uint32_t pad3;
};

struct FPhase {
    uint32_t op_begin, op_end;
    uint8_t rpos[kRegBits];   // local positions of the register bits, ascending
    uint16_t roff[kRegs];     // local index offset of register i
};

struct FPass {
    uint32_t t;               // tile qubits
    uint32_t nthreads;        // 2^(t - 4)
    uint32_t phase_begin, phase_end;
    uint32_t direct_first;    // 1: the first phase's registers are local bits t-4..t-1
    uint32_t direct_last;     // 1: same for the last phase
    uint8_t tpos[kMaxTile];   // global positions of the local bits, ascending
    uint64_t hoff[kRegs];     // global offset of local bits t-4..t-1 = j
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
// dropped.
void plan(const qvnt_gate* g, size_t n, unsigned ns, unsigned tq, std::vector<Step>& steps);

}  // namespace fusion
}  // namespace qvnt
