// Explicit axis operators for the fused 2D route (operator_plan.hpp).
#include "operator_plan.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

namespace dsmvc::hip {
namespace {

// Relative size below which the unit-vector solves stop following the decay
// of G^-1: far below the truncation tolerance, so it changes no kept entry
// by more than rounding.
constexpr double kSolveTail = 1.0 / 4503599627370496.0;   // 2^-52

struct Factors {
    int n = 0, w = 0;
    std::vector<double> d;     // [n] pivots D_i + DBL_EPSILON
    std::vector<double> l;     // [(k - 1) * n + i] = L(i + k, i), k = 1..w
    [[nodiscard]] double lower(int k, int i) const { return l[static_cast<std::size_t>(k - 1) * n + i]; }
};

// G = A^T A from the CSR of A^T (rows sorted by source index), banded with
// half bandwidth w, then G = L D L^T in FP64; the pivots carry the same
// +DBL_EPSILON as the planner's factors.
Factors factor(const AxisPlan &plan, const std::vector<double> &wt) {
    Factors f;
    f.n = plan.destination_size;
    f.w = plan.half_bandwidth;
    const int n = f.n, w = f.w;
    const auto &off = plan.transpose_offsets;
    const auto &idx = plan.transpose_indices;
    std::vector<double> g(static_cast<std::size_t>(w + 1) * n, 0.0);   // [k * n + i] = G(i, i + k)
    for (int i = 0; i < n; ++i)
        for (int k = 0; k <= w && i + k < n; ++k) {
            double s = 0.0;
            std::uint32_t a = off[i], b = off[i + k];
            const std::uint32_t ae = off[i + 1], be = off[i + k + 1];
            while (a < ae && b < be) {
                if (idx[a] < idx[b]) ++a;
                else if (idx[b] < idx[a]) ++b;
                else s += wt[a++] * wt[b++];
            }
            g[static_cast<std::size_t>(k) * n + i] = s;
        }
    f.d.assign(n, 0.0);
    f.l.assign(static_cast<std::size_t>(std::max(w, 1)) * n, 0.0);
    std::vector<double> dd(n, 0.0);   // unshifted pivots
    for (int i = 0; i < n; ++i) {
        double di = g[i];
        for (int k = 1; k <= w && i - k >= 0; ++k) {
            const double lik = f.lower(k, i - k);
            di -= lik * lik * dd[i - k];
        }
        dd[i] = di;
        f.d[i] = di + DBL_EPSILON;
        for (int k = 1; k <= w && i + k < n; ++k) {
            // L(i + k, i) = (G(i, i + k) - sum_j L(i + k, i - j) L(i, i - j) D_{i-j}) / D_i
            double s = g[static_cast<std::size_t>(k) * n + i];
            for (int j = 1; j + k <= w && i - j >= 0; ++j)
                s -= f.lower(k + j, i - j) * f.lower(j, i - j) * dd[i - j];
            f.l[static_cast<std::size_t>(k - 1) * n + i] = s / f.d[i];
        }
    }
    return f;
}

std::uint16_t half_bits(double v) {
    const _Float16 h = static_cast<_Float16>(static_cast<float>(v));
    std::uint16_t b;
    std::memcpy(&b, &h, sizeof b);
    return b;
}
double half_value(std::uint16_t b) {
    _Float16 h;
    std::memcpy(&h, &b, sizeof h);
    return static_cast<double>(static_cast<float>(h));
}
// bfloat16 by round-to-nearest-even of the F32 rounding.
std::uint16_t bf16_bits(double v) {
    const float f = static_cast<float>(v);
    std::uint32_t u;
    std::memcpy(&u, &f, sizeof u);
    u += 0x7FFFU + ((u >> 16) & 1U);
    return static_cast<std::uint16_t>(u >> 16);
}
double bf16_value(std::uint16_t b) {
    const std::uint32_t u = static_cast<std::uint32_t>(b) << 16;
    float f;
    std::memcpy(&f, &u, sizeof f);
    return static_cast<double>(f);
}

} // namespace

bool build_operator_plan(const AxisPlan &plan, OperatorPlan &out) {
    out = OperatorPlan{};
    const int n = plan.destination_size, m = plan.source_size, w = plan.half_bandwidth;
    if (n < 1 || m < 1 || w < 0) return false;
    out.n = n;
    out.m = m;
    const auto &off = plan.transpose_offsets;
    const auto &idx = plan.transpose_indices;
    std::vector<double> wt(plan.transpose_weights.size());
    if (plan.transpose_weights_f64.size() == wt.size())
        wt = plan.transpose_weights_f64;
    else
        for (std::size_t k = 0; k < wt.size(); ++k) wt[k] = plan.transpose_weights[k];
    const Factors f = factor(plan, wt);

    out.lo.assign(n, 0);
    out.count.assign(n, 0);
    out.start.assign(static_cast<std::size_t>(n) + 1, 0);
    std::vector<double> u(n, 0.0), g(n, 0.0), row;
    for (int j = 0; j < n; ++j) {
        // L u = e_j: zero before j, decaying after it.
        std::fill(u.begin(), u.end(), 0.0);
        u[j] = 1.0;
        int hi = j;
        for (int i = j + 1; i < n; ++i) {
            double s = 0.0;
            for (int k = 1; k <= w && i - k >= j; ++k) s -= f.lower(k, i - k) * u[i - k];
            u[i] = s;
            hi = i;
            if (i - j >= w) {
                bool small = true;
                for (int k = 0; k < std::max(w, 1) && small; ++k) small = std::fabs(u[i - k]) < kSolveTail;
                if (small) break;
            }
        }
        for (int i = j; i <= hi; ++i) u[i] /= f.d[i];
        // L^T g = D^-1 u, from hi down: decaying below j.
        double gmax = 0.0;
        int lo = 0;
        for (int i = hi; i >= 0; --i) {
            double s = i >= j ? u[i] : 0.0;
            for (int k = 1; k <= w && i + k <= hi; ++k) s -= f.lower(k, i) * g[i + k];
            g[i] = s;
            gmax = std::max(gmax, std::fabs(s));
            lo = i;
            if (i < j - w) {
                bool small = true;
                for (int k = 0; k < std::max(w, 1) && small; ++k)
                    small = std::fabs(g[i + k]) < kSolveTail * gmax;
                if (small) break;
            }
        }
        // M(j, c) = sum_i g_i A^T(i, c) over rows lo..hi of A^T.
        int cmin = m, cmax = -1;
        for (int i = lo; i <= hi; ++i)
            if (off[i] < off[i + 1]) {
                cmin = std::min(cmin, idx[off[i]]);
                cmax = std::max(cmax, idx[off[i + 1] - 1]);
            }
        if (cmax < cmin) {   // an all-zero row (no taps)
            out.lo[j] = 0;
            out.count[j] = 0;
            out.start[j + 1] = static_cast<std::uint32_t>(out.values.size());
            continue;
        }
        row.assign(static_cast<std::size_t>(cmax - cmin + 1), 0.0);
        for (int i = lo; i <= hi; ++i)
            for (std::uint32_t k = off[i]; k < off[i + 1]; ++k) row[idx[k] - cmin] += g[i] * wt[k];
        double total = 0.0;
        for (double v : row) total += std::fabs(v);
        const double allow = kOperatorTolerance * total;
        std::size_t a = 0, b = row.size();
        double cut = 0.0;
        while (a < b && cut + std::fabs(row[a]) <= allow) cut += std::fabs(row[a++]);
        double dropped = cut;
        cut = 0.0;
        while (b > a && cut + std::fabs(row[b - 1]) <= allow) cut += std::fabs(row[--b]);
        dropped = std::max(dropped, cut);
        if (total > 0.0) out.dropped = std::max(out.dropped, dropped / total);
        out.lo[j] = cmin + static_cast<int>(a);
        out.count[j] = static_cast<int>(b - a);
        double l1 = 0.0;
        for (std::size_t k = a; k < b; ++k) {
            out.values.push_back(row[k]);
            out.max_abs = std::max(out.max_abs, std::fabs(row[k]));
            l1 += std::fabs(row[k]);
        }
        out.max_row_l1 = std::max(out.max_row_l1, l1);
        out.start[j + 1] = static_cast<std::uint32_t>(out.values.size());
    }

    // Blocks of kOpBlock destination indices, padded to whole tiles; a
    // padding block repeats the last real block's K range with zeros.
    const int real = (n + kOpBlock - 1) / kOpBlock;
    out.blocks = (real + kOpTileBlocks - 1) / kOpTileBlocks * kOpTileBlocks;
    out.kstart.assign(out.blocks, 0);
    int steps = 1;
    for (int b = 0; b < real; ++b) {
        int klo = m, khi = 0;
        for (int j = b * kOpBlock; j < std::min(n, (b + 1) * kOpBlock); ++j)
            if (out.count[j] > 0) {
                klo = std::min(klo, out.lo[j]);
                khi = std::max(khi, out.lo[j] + out.count[j]);
            }
        if (khi <= klo) klo = khi = (b > 0 ? out.kstart[b - 1] : 0);
        out.kstart[b] = klo / 8 * 8;
        steps = std::max(steps, (khi - out.kstart[b] + kOpStep - 1) / kOpStep);
    }
    for (int b = real; b < out.blocks; ++b) out.kstart[b] = out.kstart[real - 1];
    out.steps = steps;
    bool monotone = true;
    for (int b = 1; b < out.blocks; ++b) monotone = monotone && out.kstart[b] >= out.kstart[b - 1];
    out.monotone = monotone;
    out.h_feasible = operator_fits(out, true, steps);
    out.v_feasible = operator_fits(out, false, steps);
    return out.h_feasible || out.v_feasible;
}

bool operator_fits(const OperatorPlan &op, bool rows, int steps) {
    if (!op.monotone || steps < op.steps || steps > kOpMaxSteps || op.max_abs > 16384.0
        || op.dropped > kOperatorTolerance)
        return false;
    if (rows) {
        for (int t = 0; t < op.blocks / kOpTileBlocks; ++t)
            if (op.kstart[4 * t + 3] - op.kstart[4 * t] + kOpStep * steps > kOpStageCols) return false;
    } else {
        const int units = op.blocks / 2;
        for (int u = 0; u < units; ++u)
            if (op.kstart[2 * u + 1] - op.kstart[2 * u] + kOpStep * steps + kOpSlab - 1 > kOpRing)
                return false;
        // A unit completes with the slab that holds its last block's K
        // start; slabs are aligned to kOpSlab rows.  At most kOpMaxUnitsPerSlab
        // of the units the kernel runs (ceil(n / kOpUnit)) may complete with
        // one slab.
        const int run = (op.n + kOpUnit - 1) / kOpUnit;
        const auto slab_of = [&](int u) { return (op.kstart[2 * u + 1] + kOpSlab - 1) / kOpSlab; };
        for (int u = 0; u + kOpMaxUnitsPerSlab < run; ++u)
            if (slab_of(u + kOpMaxUnitsPerSlab) == slab_of(u)) return false;
    }
    return true;
}

std::vector<std::uint32_t> pack_operator(const OperatorPlan &op, bool bf16, int steps) {
    const int parts = bf16 ? 3 : 2;
    std::vector<std::uint32_t> data(static_cast<std::size_t>(op.blocks) * steps * parts * 64 * 4, 0U);
    for (int b = 0; b < op.blocks; ++b)
        for (int s = 0; s < steps; ++s)
            for (int l = 0; l < 64; ++l) {
                const int j = b * kOpBlock + l % 16;
                std::uint16_t bits[3][8] = {};
                for (int e = 0; e < 8; ++e) {
                    const int c = op.kstart[b] + s * kOpStep + 8 * (l / 16) + e;
                    double v = 0.0;
                    if (j < op.n && c >= op.lo[j] && c < op.lo[j] + op.count[j])
                        v = op.values[op.start[j] + static_cast<std::uint32_t>(c - op.lo[j])];
                    for (int p = 0; p < parts; ++p) {
                        bits[p][e] = bf16 ? bf16_bits(v) : half_bits(v);
                        v -= bf16 ? bf16_value(bits[p][e]) : half_value(bits[p][e]);
                    }
                }
                for (int p = 0; p < parts; ++p) {
                    std::uint32_t *dst = &data[((((static_cast<std::size_t>(b) * steps + s) * parts + p) * 64
                                                 + l) * 4)];
                    for (int q = 0; q < 4; ++q)
                        dst[q] = static_cast<std::uint32_t>(bits[p][2 * q])
                                 | (static_cast<std::uint32_t>(bits[p][2 * q + 1]) << 16);
                }
            }
    return data;
}

} // namespace dsmvc::hip
