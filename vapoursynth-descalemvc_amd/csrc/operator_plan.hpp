// Explicit operator of one axis plan for the fused 2D route: the rows of
// M = G^-1 A^T (G = A^T A), truncated to the band that carries all but a
// 2^-27 share of each row's L1 mass, blocked and packed as MFMA operands.
// Pure host code (FP64), unit-testable without a device.
//
// The descale of one axis is x = G^-1 A^T y, which the recurrence kernels
// compute as a gather followed by the banded LDL^T substitutions.  M is the
// same linear map written out: G^-1 decays geometrically away from the
// diagonal (the decay the windowed partition relies on), so every row of M
// is a short run of source indices around the destination's position, and a
// 2D plane becomes X = M_v Y M_h^T, two banded matrix products that the
// fused kernel evaluates tile by tile without writing the horizontal result
// to memory.
#pragma once

#include <cstdint>
#include <vector>

#include "device_types.hpp"
#include "engine.hpp"

namespace dsmvc::hip {

struct OperatorPlan {
    std::int32_t n = 0, m = 0;
    std::int32_t blocks = 0;         // padded: a multiple of kOpTileBlocks
    std::int32_t steps = 0;          // K steps of kOpStep source indices per block
    double dropped = 0.0;            // largest L1 share dropped from a row
    double max_abs = 0.0;            // largest |M(j, c)|
    double max_row_l1 = 0.0;         // largest sum_c |M(j, c)| over the kept band
    std::vector<std::int32_t> lo;    // [n] first source index of row j's band
    std::vector<std::int32_t> count; // [n] band length
    std::vector<std::uint32_t> start;// [n + 1] row j's values at values[start[j] ..]
    std::vector<double> values;
    std::vector<std::int32_t> kstart;// [blocks] first source index of block b's K range
    bool monotone = false;           // kstart never decreases
    bool h_feasible = false;         // fits the horizontal staging window
    bool v_feasible = false;         // fits the vertical ring
};

// Share of a row's L1 mass that truncation may drop from each end.
constexpr double kOperatorTolerance = 1.0 / 268435456.0;

// Builds M for `plan` in FP64 from its weights (the FP64 weights when the
// plan retains them, else the F32 weights it applies) and an FP64 LDL^T of
// G.  False when the band does not fit the fused kernel in either direction
// (it then never takes the route).
bool build_operator_plan(const AxisPlan &plan, OperatorPlan &out);

// True when the fused kernel can use `op` in the given direction with every
// block's K range padded to `steps` steps (the staging window for rows, the
// ring for columns).
bool operator_fits(const OperatorPlan &op, bool rows, int steps);

// MFMA operand images: [blocks][steps][parts][64 lanes][4 words], lane l of
// block b and step s holding M(16b + l % 16, kstart[b] + 32s + 8(l / 16) + e)
// for e = 0..7 as 16-bit values.  f16: two parts (the FP16 rounding and the
// FP16 rounding of the remainder); bf16: three parts likewise.  `steps`
// (at least op.steps) pads every block's K range with zeros, so that two
// operators of different band widths share one kernel instance.
std::vector<std::uint32_t> pack_operator(const OperatorPlan &op, bool bf16, int steps);

} // namespace dsmvc::hip
