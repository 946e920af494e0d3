// Packing of axis plans into device images and chunk partitions.
#include "device_plan.hpp"
#include "trace.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "kernels.hpp"

namespace dsmvc::hip {
namespace {

void check(hipError_t e, const char *what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string("dsmvc: ") + what + ": " + hipGetErrorString(e));
}

std::size_t align256(std::size_t v) { return (v + 255U) & ~std::size_t{255}; }

// Host-to-device copy of a table, on a non-blocking stream of the current
// device, waited for before returning.  A plain hipMemcpy runs on the null
// stream, which first waits for all work queued on blocking streams; a plan
// built while the executor's streams are busy (a served request with a new
// geometry) would stall behind every queued kernel.
hipError_t upload(void *dst, const void *src, std::size_t bytes) {
    static std::mutex lock;
    static std::map<int, hipStream_t> streams;   // one per device, kept until exit
    int device = 0;
    if (const hipError_t e = hipGetDevice(&device); e != hipSuccess) return e;
    hipStream_t stream = nullptr;
    {
        std::lock_guard guard(lock);
        auto it = streams.find(device);
        if (it == streams.end()) {
            if (const hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
                e != hipSuccess)
                return e;
            streams.emplace(device, stream);
        } else {
            stream = it->second;
        }
    }
    if (const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream);
        e != hipSuccess)
        return e;
    return hipStreamSynchronize(stream);
}

} // namespace

PackedPlan pack_plan(const AxisPlan &p) {
    PackedPlan out;
    out.n = p.destination_size;
    out.m = p.source_size;
    out.w = p.half_bandwidth;
    out.nnz = static_cast<std::int32_t>(p.transpose_indices.size());
    if (out.w > kMaxWideBand)
        throw std::invalid_argument("dsmvc: half bandwidth exceeds the HIP executor's limit of 15");
    out.coef_stride = coef_stride(out.w);
    const int slots = (out.coef_stride - 2) / 2;
    const auto n = static_cast<std::size_t>(out.n);
    const auto cs = static_cast<std::size_t>(out.coef_stride);
    out.coef.assign(n * cs, 0.0F);
    out.coef64.assign(n * cs, 0.0);
    out.zform = p.requires_float64();
    for (std::size_t i = 0; i < n; ++i) {
        float *row = out.coef.data() + i * cs;
        double *row64 = out.coef64.data() + i * cs;
        row[slots] = p.inverse_diagonal[i];
        for (int d = 1; d <= out.w; ++d) {
            const std::size_t at = static_cast<std::size_t>(d - 1) * n + i;
            row[d - 1] = p.lower_ld[at];
            row[slots + d] = p.upper_l[at];
        }
        if (out.zform) {
            // Retained FP64 factors: unit-L forward rows, diagonal applied after
            // (reference src/axis_plan.cpp:1124-1191, retained branch).
            row64[slots] = p.inverse_diagonal_f64[i];
            for (int d = 1; d <= out.w; ++d) {
                if (static_cast<std::size_t>(d) <= i)
                    row64[d - 1] = p.ldlt_bands_f64[static_cast<std::size_t>(d) * n + i - static_cast<std::size_t>(d)];
                if (i + static_cast<std::size_t>(d) < n)
                    row64[slots + d] = p.ldlt_bands_f64[static_cast<std::size_t>(d) * n + i];
            }
        } else {
            for (int k = 0; k < out.coef_stride; ++k) row64[k] = static_cast<double>(row[k]);
        }
    }
    out.wt64.resize(static_cast<std::size_t>(out.nnz));
    for (std::size_t k = 0; k < out.wt64.size(); ++k)
        out.wt64[k] = out.zform ? p.transpose_weights_f64[k] : static_cast<double>(p.transpose_weights[k]);
    return out;
}

bool build_chunk_plan(const AxisPlan &p, int max_chunks, ChunkPlan &out) {
    const int n = p.destination_size, w = p.half_bandwidth;
    if (!chunked_band_supported(w) || max_chunks < 1 || max_chunks > kMaxChunks) return false;
    const int length = chunked_length_for((n + max_chunks - 1) / max_chunks);
    if (length == 0 || length < w) return false;
    int taps = 0;
    for (int i = 0; i < n; ++i)
        taps = std::max(taps, static_cast<int>(p.transpose_offsets[static_cast<std::size_t>(i) + 1]
                                               - p.transpose_offsets[static_cast<std::size_t>(i)]));
    const int width = chunked_width_for(std::max(taps, 1));
    if (width == 0) return false;
    const int count = (n + length - 1) / length;
    const int rows = count * length;
    out.chunks = count;
    out.length = length;
    out.width = width;
    const auto un = static_cast<std::size_t>(n);
    const auto ur = static_cast<std::size_t>(rows);

    // Coefficient rows, zero past n.
    out.coef.assign(ur * 16, 0.0F);
    for (std::size_t i = 0; i < un; ++i) {
        float *row = out.coef.data() + i * 16;
        row[kMaxBand] = p.inverse_diagonal[i];
        for (int d = 1; d <= w; ++d) {
            const std::size_t at = static_cast<std::size_t>(d - 1) * un + i;
            row[d - 1] = p.lower_ld[at];
            row[kMaxBand + d] = p.upper_l[at];
        }
    }

    // Staging windows of the horizontal kernel's row blocks: 16 rows when
    // every block's taps fit 32 columns, else 8.
    auto staging = [&](int g) {
        const int nb = length / g;
        std::vector<std::int32_t> lo_table(static_cast<std::size_t>(count) * nb, 0);
        bool fits = true;
        for (int q = 0; q < count; ++q) {
            int previous = 0;
            for (int b = 0; b < nb; ++b) {
                int lo = std::numeric_limits<int>::max(), hi = -1;
                for (int t = 0; t < g; ++t) {
                    const int i = q * length + b * g + t;
                    if (i >= n) break;
                    const auto k0 = p.transpose_offsets[static_cast<std::size_t>(i)];
                    const auto k1 = p.transpose_offsets[static_cast<std::size_t>(i) + 1];
                    if (k1 == k0) continue;
                    lo = std::min(lo, p.transpose_indices[k0]);
                    hi = std::max(hi, p.transpose_indices[k1 - 1]);
                }
                if (hi < 0) lo = previous;
                else if (hi - lo >= kStageCols) fits = false;
                lo_table[static_cast<std::size_t>(q) * nb + b] = lo;
                previous = lo;
            }
        }
        if (fits) {
            out.block_rows = g;
            out.blocks = nb;
            out.blk_lo = std::move(lo_table);
        }
        return fits;
    };
    out.rows_feasible = staging(16) || staging(kBlockRowsMin);
    if (!out.rows_feasible) {
        out.block_rows = kBlockRowsMin;
        out.blocks = length / kBlockRowsMin;
        out.blk_lo.assign(static_cast<std::size_t>(count) * out.blocks, 0);
    }

    // Taps packed to `width` per row, short rows padded at the front with
    // weight 0 at an index inside the row's staging window.
    out.ell_idx.assign(ur * width, 0);
    out.ell_w.assign(ur * width, 0.0F);
    for (int i = 0; i < rows; ++i) {
        const int q = i / length, b = (i % length) / out.block_rows;
        const int window = out.blk_lo[static_cast<std::size_t>(q) * out.blocks + b];
        std::int32_t *idx = out.ell_idx.data() + static_cast<std::size_t>(i) * width;
        float *wt = out.ell_w.data() + static_cast<std::size_t>(i) * width;
        if (i >= n) {
            std::fill(idx, idx + width, window);
            continue;
        }
        const auto k0 = p.transpose_offsets[static_cast<std::size_t>(i)];
        const auto k1 = p.transpose_offsets[static_cast<std::size_t>(i) + 1];
        const int c = static_cast<int>(k1 - k0);
        const int pad = width - c;
        const std::int32_t fill = c > 0 ? p.transpose_indices[k0] : window;
        for (int e = 0; e < pad; ++e) idx[e] = fill;
        for (int e = 0; e < c; ++e) {
            idx[pad + e] = p.transpose_indices[k0 + static_cast<std::uint32_t>(e)];
            wt[pad + e] = p.transpose_weights[k0 + static_cast<std::uint32_t>(e)];
        }
    }

    // Transfer matrices over the padded rows, in FP64.
    out.tf.assign(static_cast<std::size_t>(count) * w * w, 0.0F);
    out.tb.assign(static_cast<std::size_t>(count) * w * w, 0.0F);
    std::vector<double> ring(static_cast<std::size_t>(w));
    auto push = [&](double v) {
        for (int d = w - 1; d > 0; --d) ring[static_cast<std::size_t>(d)] = ring[static_cast<std::size_t>(d - 1)];
        ring[0] = v;
    };
    auto coef = [&](int i, int slot) { return static_cast<double>(out.coef[static_cast<std::size_t>(i) * 16 + slot]); };
    for (int q = 0; q < count; ++q) {
        const int a = q * length, b = a + length;
        for (int c = 0; c < w; ++c) {
            std::fill(ring.begin(), ring.end(), 0.0);
            ring[static_cast<std::size_t>(c)] = 1.0;
            for (int i = a; i < b; ++i) {
                double v = 0.0;
                for (int d = w; d >= 1; --d) v -= coef(i, d - 1) * ring[static_cast<std::size_t>(d - 1)];
                push(v * coef(i, kMaxBand));
            }
            for (int r = 0; r < w; ++r)
                out.tf[(static_cast<std::size_t>(q) * w + r) * w + c] = static_cast<float>(ring[static_cast<std::size_t>(r)]);
            std::fill(ring.begin(), ring.end(), 0.0);
            ring[static_cast<std::size_t>(c)] = 1.0;
            for (int i = b - 1; i >= a; --i) {
                double v = 0.0;
                for (int d = w; d >= 1; --d) v -= coef(i, kMaxBand + d) * ring[static_cast<std::size_t>(d - 1)];
                push(v);
            }
            for (int r = 0; r < w; ++r)
                out.tb[(static_cast<std::size_t>(q) * w + r) * w + c] = static_cast<float>(ring[static_cast<std::size_t>(r)]);
        }
    }
    return true;
}

namespace {

// A windowed partition of either precision for `length`-row chunks with
// `warm` warm-up rows.  `coef_of(i, slot)` is plan row i's coefficient record
// in the kMaxBand layout and `weight_of(q)` CSR weight q.  `zform`: the
// forward recurrence carries z = L^-1 b and the diagonal multiplies only the
// stored value (retained FP64 factors).
template <class T, class Coef, class Weight>
bool build_window(const AxisPlan &p, int length, int warm, Coef coef_of, Weight weight_of,
                  bool zform, double tolerance, WindowPlanT<T> &out) {
    const int n = p.destination_size, w = p.half_bandwidth;
    const int band = window_band_class(w);
    if (band == 0 || n < 1) return false;
    int taps = 0;
    for (int i = 0; i < n; ++i)
        taps = std::max(taps, static_cast<int>(p.transpose_offsets[static_cast<std::size_t>(i) + 1]
                                               - p.transpose_offsets[static_cast<std::size_t>(i)]));
    // The FP64 kernels take tap widths 8 and 16, the F32 kernels also 6.
    const int width = std::is_same_v<T, float> ? window_width_for(band, std::max(taps, 1))
                                               : chunked_width_for(std::max(taps, 1));
    if (width == 0) return false;
    const int chunks = (n + length - 1) / length;
    const int rows = chunks * length + 2 * warm;
    out.band = band;
    out.chunks = chunks;
    out.length = length;
    out.warmup = warm;
    out.width = width;
    const auto ur = static_cast<std::size_t>(rows);
    auto real = [&](int r) { return r - warm; };   // plan row of padded row r

    out.zform = zform;
    out.coef.assign(ur * 16, T(0));
    for (int r = 0; r < rows; ++r) {
        const int i = real(r);
        if (i < 0 || i >= n) continue;
        T *row = out.coef.data() + static_cast<std::size_t>(r) * 16;
        for (int slot = 0; slot < 16; ++slot) row[slot] = coef_of(i, slot);
    }

    // Staging windows of 8- and 16-row blocks; blocks of zero rows reuse a
    // neighbouring window (their taps carry weight 0).
    auto staging = [&](int g, std::vector<std::int32_t> &table) {
        const int nb = rows / g;
        table.assign(static_cast<std::size_t>(nb), 0);
        bool fits = true;
        int previous = 0;
        for (int b = 0; b < nb; ++b) {
            int lo = std::numeric_limits<int>::max(), hi = -1;
            for (int t = 0; t < g; ++t) {
                const int i = real(b * g + t);
                if (i < 0 || i >= n) continue;
                const auto k0 = p.transpose_offsets[static_cast<std::size_t>(i)];
                const auto k1 = p.transpose_offsets[static_cast<std::size_t>(i) + 1];
                if (k1 == k0) continue;
                lo = std::min(lo, p.transpose_indices[k0]);
                hi = std::max(hi, p.transpose_indices[k1 - 1]);
            }
            if (hi < 0) lo = previous;
            lo &= ~1;   // even: integer samples are staged in aligned pairs
            if (hi >= 0 && hi - lo >= kStageCols) fits = false;
            table[static_cast<std::size_t>(b)] = lo;
            previous = lo;
        }
        return fits;
    };
    out.rows_feasible = staging(kWindowBlock, out.blk_lo);
    out.block16 = staging(2 * kWindowBlock, out.blk_lo16);

    // Contiguous tap runs for the horizontal kernel: row i reads columns
    // [s, s + K) of its staging window, s = max(lo8, last - K + 1), with its
    // CSR weights at their columns and 0 elsewhere.  The run must hold the
    // whole row and lie inside the 8-row window (and the 16-row window when
    // those are used); otherwise the horizontal pass is not windowed.
    out.tap_start.assign(ur, 0);
    out.tap_w.assign(ur * width, T(0));
    for (int r = 0; r < rows; ++r) {
        const int i = real(r);
        const int lo8 = out.blk_lo[static_cast<std::size_t>(r / kWindowBlock)];
        const int lo16 = out.blk_lo16[static_cast<std::size_t>(r / (2 * kWindowBlock))];
        int start = lo8;
        if (i >= 0 && i < n) {
            const auto k0 = p.transpose_offsets[static_cast<std::size_t>(i)];
            const auto k1 = p.transpose_offsets[static_cast<std::size_t>(i) + 1];
            if (k1 > k0) {
                const int first = p.transpose_indices[k0], last = p.transpose_indices[k1 - 1];
                start = std::max(lo8, last - width + 1);
                if (last - first + 1 > width) out.rows_feasible = false;
                T *wt = out.tap_w.data() + static_cast<std::size_t>(r) * width;
                for (auto q = k0; q < k1; ++q) {
                    const int e = p.transpose_indices[q] - start;
                    if (e >= 0 && e < width) wt[e] = weight_of(q);
                }
            }
        }
        if (start + width > lo8 + kStageCols) out.rows_feasible = false;
        if (start < lo16 || start + width > lo16 + kStageCols) out.block16 = false;
        out.tap_start[static_cast<std::size_t>(r)] = start;
    }

    // Packed records: forward [tap weights | lld_1..lld_W | inverse diagonal],
    // backward [ul_1..ul_W], zero in padding rows and past the plan's band.
    const int fw = window_fwd_words(band, width), bw = window_bwd_words(band);
    out.fwd.assign(ur * static_cast<std::size_t>(fw), T(0));
    out.bwd.assign(ur * static_cast<std::size_t>(bw), T(0));
    for (int r = 0; r < rows; ++r) {
        const T *c = out.coef.data() + static_cast<std::size_t>(r) * 16;
        T *f = out.fwd.data() + static_cast<std::size_t>(r) * fw;
        T *u = out.bwd.data() + static_cast<std::size_t>(r) * bw;
        std::copy_n(out.tap_w.data() + static_cast<std::size_t>(r) * width, width, f);
        for (int d = 1; d <= w; ++d) {
            f[width + d - 1] = c[d - 1];
            u[d - 1] = c[kMaxBand + d];
        }
        f[width + band] = c[kMaxBand];
    }

    // Warm-up check, in FP64 on the coefficients the kernels use: from
    // each unit starting state P rows before a chunk boundary (forward) or
    // after it (backward), the state must stay below the tolerance over the
    // whole chunk.  Boundaries next to the plan's ends start in zero rows and
    // are exact.
    auto coef = [&](int r, int slot) {
        return static_cast<double>(out.coef[static_cast<std::size_t>(r) * 16 + static_cast<std::size_t>(slot)]);
    };
    // The last `band` states in a circular buffer: back(d) is the state d
    // rows back (d = 1..band), push() makes v the newest.
    std::array<double, kMaxBand> ring{};
    int head = 0;
    auto back = [&](int d) {
        int at = head + d - 1;
        if (at >= band) at -= band;
        return ring[static_cast<std::size_t>(at)];
    };
    auto push = [&](double v) {
        head = head == 0 ? band - 1 : head - 1;
        ring[static_cast<std::size_t>(head)] = v;
    };
    auto reset = [&](int c) {
        ring.fill(0.0);
        head = 0;
        ring[static_cast<std::size_t>(c)] = 1.0;
    };
    auto peak = [&] {
        double m = 0.0;
        for (int k = 0; k < band; ++k) m = std::max(m, std::fabs(ring[static_cast<std::size_t>(k)]));
        return m;
    };
    double worst = 0.0;
    for (int q = 1; q < chunks; ++q) {
        const int a = q * length + warm;      // padded row of the chunk's first row
        for (int c = 0; c < band; ++c) {
            reset(c);
            for (int r = a - warm; r < a + length; ++r) {
                double v = 0.0;
                for (int d = band; d >= 1; --d) v -= coef(r, d - 1) * back(d);
                push(zform ? v : v * coef(r, kMaxBand));
                if (r >= a - 1) worst = std::max(worst, peak());
            }
        }
    }
    for (int q = 0; q + 1 < chunks; ++q) {
        const int b = (q + 1) * length + warm;   // padded row just past the chunk
        for (int c = 0; c < band; ++c) {
            reset(c);
            for (int r = b + warm - 1; r >= b - length; --r) {
                double v = 0.0;
                for (int d = band; d >= 1; --d) v -= coef(r, kMaxBand + d) * back(d);
                push(v);
                if (r <= b) worst = std::max(worst, peak());
            }
        }
    }
    out.residual = worst;
    // Both passes stage 32 source samples per block, so the runs must fit.
    return out.rows_feasible && worst <= tolerance;
}

} // namespace

bool build_window_plan(const AxisPlan &p, WindowPlan &out) {
    const int band = window_band_class(p.half_bandwidth);
    if (band == 0) return false;
    const auto un = static_cast<std::size_t>(p.destination_size);
    const int w = p.half_bandwidth;
    auto coef = [&](int i, int slot) -> float {
        const auto ui = static_cast<std::size_t>(i);
        if (slot == kMaxBand) return p.inverse_diagonal[ui];
        const int d = slot < kMaxBand ? slot + 1 : slot - kMaxBand;
        if (d > w || slot > 2 * kMaxBand) return 0.0F;
        const std::size_t at = static_cast<std::size_t>(d - 1) * un + ui;
        return slot < kMaxBand ? p.lower_ld[at] : p.upper_l[at];
    };
    auto weight = [&](std::uint32_t q) { return p.transpose_weights[q]; };
    return build_window<float>(p, window_length(band), window_warmup(band), coef, weight, false,
                               kWindowTolerance, out);
}

bool build_solve_window_plan(const AxisPlan &p, WindowPlan &out) {
    const int band = window_band_class(p.half_bandwidth);
    const int n = p.destination_size, w = p.half_bandwidth;
    if (band == 0 || n < 1) return false;
    // The same system with identity taps: row i reads element i of its
    // right-hand side.
    AxisPlan q;
    q.source_size = n;
    q.destination_size = n;
    q.half_bandwidth = w;
    q.transpose_offsets.resize(static_cast<std::size_t>(n) + 1);
    q.transpose_indices.resize(static_cast<std::size_t>(n));
    for (int i = 0; i <= n; ++i) q.transpose_offsets[static_cast<std::size_t>(i)] = static_cast<std::uint32_t>(i);
    for (int i = 0; i < n; ++i) q.transpose_indices[static_cast<std::size_t>(i)] = i;
    const auto un = static_cast<std::size_t>(n);
    auto coef = [&](int i, int slot) -> float {
        const auto ui = static_cast<std::size_t>(i);
        if (slot == kMaxBand) return p.inverse_diagonal[ui];
        const int d = slot < kMaxBand ? slot + 1 : slot - kMaxBand;
        if (d > w || slot > 2 * kMaxBand) return 0.0F;
        const std::size_t at = static_cast<std::size_t>(d - 1) * un + ui;
        return slot < kMaxBand ? p.lower_ld[at] : p.upper_l[at];
    };
    auto weight = [](std::uint32_t) { return 1.0F; };
    return build_window<float>(q, window_length(band), window_warmup(band), coef, weight, false,
                               kWindowTolerance, out);
}

bool build_window_plan64(const AxisPlan &p, const PackedPlan &packed, WindowPlan64 &out) {
    const int band = window_band_class(p.half_bandwidth);
    if (band == 0 || packed.coef_stride != 16) return false;
    auto coef = [&](int i, int slot) {
        return packed.coef64[static_cast<std::size_t>(i) * 16 + static_cast<std::size_t>(slot)];
    };
    auto weight = [&](std::uint32_t q) { return packed.wt64[q]; };
    return build_window<double>(p, window_length64(band), window_warmup64(band), coef, weight,
                                packed.zform != 0, kWindowTolerance64, out);
}

DevicePlan::DevicePlan(const AxisPlan &plan, int device, ByteAccount account)
    : keep_(std::make_shared<const AxisPlan>(plan)), plan_(keep_.get()), device_(device),
      account_(std::move(account)) {
    const PackedPlan packed = pack_plan(*plan_);
    const auto n = static_cast<std::size_t>(packed.n);
    const auto nnz = static_cast<std::size_t>(packed.nnz);
    const std::size_t o_off = 0;
    const std::size_t o_idx = align256(o_off + (n + 1) * 4);
    const std::size_t o_wt = align256(o_idx + std::max<std::size_t>(nnz, 1) * 4);
    const std::size_t o_coef = align256(o_wt + std::max<std::size_t>(nnz, 1) * 4);
    const std::size_t o_wt64 = align256(o_coef + packed.coef.size() * 4);
    const std::size_t o_coef64 = align256(o_wt64 + std::max<std::size_t>(nnz, 1) * 8);
    bytes_ = align256(o_coef64 + packed.coef64.size() * 8);
    std::vector<unsigned char> host(bytes_, 0);
    std::memcpy(host.data() + o_off, plan_->transpose_offsets.data(), (n + 1) * 4);
    std::memcpy(host.data() + o_idx, plan_->transpose_indices.data(), nnz * 4);
    std::memcpy(host.data() + o_wt, plan_->transpose_weights.data(), nnz * 4);
    std::memcpy(host.data() + o_coef, packed.coef.data(), packed.coef.size() * 4);
    std::memcpy(host.data() + o_wt64, packed.wt64.data(), nnz * 8);
    std::memcpy(host.data() + o_coef64, packed.coef64.data(), packed.coef64.size() * 8);
    check(hipMalloc(&memory_, bytes_), "device plan allocation failed");
    const hipError_t copy = upload(memory_, host.data(), bytes_);
    if (copy != hipSuccess) {
        (void)hipFree(memory_);
        memory_ = nullptr;
        check(copy, "device plan upload failed");
    }
    grow(bytes_);
    auto *base = static_cast<unsigned char *>(memory_);
    axis_.n = packed.n;
    axis_.m = packed.m;
    axis_.w = packed.w;
    axis_.nnz = packed.nnz;
    axis_.off = reinterpret_cast<const std::uint32_t *>(base + o_off);
    axis_.idx = reinterpret_cast<const std::int32_t *>(base + o_idx);
    axis_.wt = reinterpret_cast<const float *>(base + o_wt);
    axis_.coef = reinterpret_cast<const float *>(base + o_coef);
    axis_.wt64 = reinterpret_cast<const double *>(base + o_wt64);
    axis_.coef64 = reinterpret_cast<const double *>(base + o_coef64);
    axis_.zform = packed.zform ? 1 : 0;
}

void DevicePlan::grow(std::size_t bytes) {
    total_bytes_.fetch_add(bytes, std::memory_order_relaxed);
    if (account_) account_->fetch_add(bytes, std::memory_order_relaxed);
}

DevicePlan::~DevicePlan() {
    if (account_) account_->fetch_sub(total_bytes_.load(std::memory_order_relaxed),
                                      std::memory_order_relaxed);
    for (auto &entry : chunk_images_)
        if (entry.second.memory) (void)hipFree(entry.second.memory);
    if (window_image_.memory) (void)hipFree(window_image_.memory);
    if (window64_image_.memory) (void)hipFree(window64_image_.memory);
    for (auto &entry : op_images_)
        if (entry.second.memory) (void)hipFree(entry.second.memory);
    if (solve_image_.memory) (void)hipFree(solve_image_.memory);
    if (refine_memory_) (void)hipFree(refine_memory_);
    if (memory_) (void)hipFree(memory_);
}

const DevChunks *DevicePlan::chunks(int max_chunks, bool rows) {
    std::lock_guard lock(mutex_);
    auto it = chunk_images_.find(max_chunks);
    if (it == chunk_images_.end()) {
        ChunkImage image;
        ChunkPlan cp;
        if (build_chunk_plan(*plan_, max_chunks, cp)) {
            const std::size_t o_tf = 0;
            const std::size_t o_tb = align256(o_tf + cp.tf.size() * 4);
            const std::size_t o_idx = align256(o_tb + cp.tb.size() * 4);
            const std::size_t o_w = align256(o_idx + cp.ell_idx.size() * 4);
            const std::size_t o_coef = align256(o_w + cp.ell_w.size() * 4);
            const std::size_t o_blk = align256(o_coef + cp.coef.size() * 4);
            const std::size_t total = align256(o_blk + cp.blk_lo.size() * 4);
            std::vector<unsigned char> host(total, 0);
            std::memcpy(host.data() + o_tf, cp.tf.data(), cp.tf.size() * 4);
            std::memcpy(host.data() + o_tb, cp.tb.data(), cp.tb.size() * 4);
            std::memcpy(host.data() + o_idx, cp.ell_idx.data(), cp.ell_idx.size() * 4);
            std::memcpy(host.data() + o_w, cp.ell_w.data(), cp.ell_w.size() * 4);
            std::memcpy(host.data() + o_coef, cp.coef.data(), cp.coef.size() * 4);
            std::memcpy(host.data() + o_blk, cp.blk_lo.data(), cp.blk_lo.size() * 4);
            int current = 0;
            check(hipGetDevice(&current), "hipGetDevice failed");
            check(hipSetDevice(device_), "hipSetDevice failed");
            check(hipMalloc(&image.memory, total), "chunk table allocation failed");
            image.bytes = total;
            if (const hipError_t e = upload(image.memory, host.data(), total); e != hipSuccess) {
                (void)hipFree(image.memory);
                check(e, "chunk table upload failed");
            }
            (void)hipSetDevice(current);
            auto *b = static_cast<unsigned char *>(image.memory);
            image.dev.chunks = cp.chunks;
            image.dev.length = cp.length;
            image.dev.ell_width = cp.width;
            image.dev.blocks = cp.blocks;
            image.dev.block_rows = cp.block_rows;
            image.dev.tf = reinterpret_cast<const float *>(b + o_tf);
            image.dev.tb = reinterpret_cast<const float *>(b + o_tb);
            image.dev.ell_idx = reinterpret_cast<const std::int32_t *>(b + o_idx);
            image.dev.ell_w = reinterpret_cast<const float *>(b + o_w);
            image.dev.coef = reinterpret_cast<const float *>(b + o_coef);
            image.dev.blk_lo = reinterpret_cast<const std::int32_t *>(b + o_blk);
            image.ok = true;
            image.rows_feasible = cp.rows_feasible;
            grow(image.bytes);
        }
        it = chunk_images_.emplace(max_chunks, image).first;
    }
    if (!it->second.ok || (rows && !it->second.rows_feasible)) return nullptr;
    return &it->second.dev;
}

// Uploads a window plan's tables as one allocation and fills `dev` (its
// plan sizes from this plan).  Returns the allocation.
void *DevicePlan::upload_window(const WindowPlan &wp, DevWindow &dev, std::size_t &bytes) {
    const std::size_t o_fwd = 0;
    const std::size_t o_bwd = align256(o_fwd + wp.fwd.size() * 4);
    const std::size_t o_ts = align256(o_bwd + wp.bwd.size() * 4);
    const std::size_t o_blk = align256(o_ts + wp.tap_start.size() * 4);
    const std::size_t o_blk16 = align256(o_blk + wp.blk_lo.size() * 4);
    const std::size_t total = align256(o_blk16 + wp.blk_lo16.size() * 4);
    std::vector<unsigned char> host(total, 0);
    std::memcpy(host.data() + o_fwd, wp.fwd.data(), wp.fwd.size() * 4);
    std::memcpy(host.data() + o_bwd, wp.bwd.data(), wp.bwd.size() * 4);
    std::memcpy(host.data() + o_ts, wp.tap_start.data(), wp.tap_start.size() * 4);
    std::memcpy(host.data() + o_blk, wp.blk_lo.data(), wp.blk_lo.size() * 4);
    std::memcpy(host.data() + o_blk16, wp.blk_lo16.data(), wp.blk_lo16.size() * 4);
    int current = 0;
    check(hipGetDevice(&current), "hipGetDevice failed");
    check(hipSetDevice(device_), "hipSetDevice failed");
    void *memory = nullptr;
    check(hipMalloc(&memory, total), "window table allocation failed");
    if (const hipError_t e = upload(memory, host.data(), total);
        e != hipSuccess) {
        (void)hipFree(memory);
        check(e, "window table upload failed");
    }
    (void)hipSetDevice(current);
    bytes = total;
    auto *b = static_cast<unsigned char *>(memory);
    dev.n = plan_->destination_size;
    dev.m = plan_->source_size;
    dev.length = wp.length;
    dev.warmup = wp.warmup;
    dev.chunks = wp.chunks;
    dev.ell_width = wp.width;
    dev.block16 = wp.block16 ? 1 : 0;
    dev.fwd = reinterpret_cast<const float *>(b + o_fwd);
    dev.bwd = reinterpret_cast<const float *>(b + o_bwd);
    dev.tap_start = reinterpret_cast<const std::int32_t *>(b + o_ts);
    dev.blk_lo = reinterpret_cast<const std::int32_t *>(b + o_blk);
    dev.blk_lo16 = reinterpret_cast<const std::int32_t *>(b + o_blk16);
    return memory;
}

const DevWindow *DevicePlan::window(bool rows) {
    std::lock_guard lock(mutex_);
    WindowImage &image = window_image_;
    if (!image.built) {
        const trace::Range range(trace::Label::window_build);
        // Marked built only once the tables are built and uploaded: a failed
        // upload throws here and the next call tries again.
        WindowPlan wp;
        if (build_window_plan(*plan_, wp)) {
            image.memory = upload_window(wp, image.dev, image.bytes);
            grow(image.bytes);
            image.ok = true;
            image.rows_feasible = wp.rows_feasible;
        }
        image.built = true;
    }
    if (!image.ok || (rows && !image.rows_feasible)) return nullptr;
    return &image.dev;
}

const DevWindow *DevicePlan::built_window(bool rows) {
    std::lock_guard lock(mutex_);
    const WindowImage &image = window_image_;
    if (!image.built || !image.ok || (rows && !image.rows_feasible)) return nullptr;
    return &image.dev;
}

const DevWindow64 *DevicePlan::built_window64() {
    std::lock_guard lock(mutex_);
    return window64_image_.built && window64_image_.ok ? &window64_image_.dev : nullptr;
}

const DevWindow *DevicePlan::solve_window(bool rows) {
    std::lock_guard lock(mutex_);
    WindowImage &image = solve_image_;
    if (!image.built) {
        const trace::Range range(trace::Label::window_build);
        WindowPlan wp;
        if (build_solve_window_plan(*plan_, wp)) {
            image.memory = upload_window(wp, image.dev, image.bytes);
            grow(image.bytes);
            image.dev.m = plan_->destination_size;
            image.ok = true;
            image.rows_feasible = wp.rows_feasible;
        }
        image.built = true;
    }
    if (!image.ok || (rows && !image.rows_feasible)) return nullptr;
    return &image.dev;
}

const DevWindow64 *DevicePlan::window64() {
    std::lock_guard lock(mutex_);
    Window64Image &image = window64_image_;
    if (!image.built) {
        const trace::Range range(trace::Label::window_build);
        WindowPlan64 wp;
        if (build_window_plan64(*plan_, pack_plan(*plan_), wp)) {
            const std::size_t o_fwd = 0;
            const std::size_t o_bwd = align256(o_fwd + wp.fwd.size() * 8);
            const std::size_t o_ts = align256(o_bwd + wp.bwd.size() * 8);
            const std::size_t o_blk = align256(o_ts + wp.tap_start.size() * 4);
            const std::size_t total = align256(o_blk + wp.blk_lo.size() * 4);
            std::vector<unsigned char> host(total, 0);
            std::memcpy(host.data() + o_fwd, wp.fwd.data(), wp.fwd.size() * 8);
            std::memcpy(host.data() + o_bwd, wp.bwd.data(), wp.bwd.size() * 8);
            std::memcpy(host.data() + o_ts, wp.tap_start.data(), wp.tap_start.size() * 4);
            std::memcpy(host.data() + o_blk, wp.blk_lo.data(), wp.blk_lo.size() * 4);
            int current = 0;
            check(hipGetDevice(&current), "hipGetDevice failed");
            check(hipSetDevice(device_), "hipSetDevice failed");
            void *memory = nullptr;
            check(hipMalloc(&memory, total), "window table allocation failed");
            if (const hipError_t e = upload(memory, host.data(), total);
                e != hipSuccess) {
                (void)hipFree(memory);
                check(e, "window table upload failed");
            }
            (void)hipSetDevice(current);
            image.memory = memory;
            image.bytes = total;
            grow(total);
            auto *b = static_cast<unsigned char *>(image.memory);
            image.dev.n = plan_->destination_size;
            image.dev.m = plan_->source_size;
            image.dev.length = wp.length;
            image.dev.warmup = wp.warmup;
            image.dev.chunks = wp.chunks;
            image.dev.ell_width = wp.width;
            image.dev.zform = wp.zform ? 1 : 0;
            image.dev.fwd = reinterpret_cast<const double *>(b + o_fwd);
            image.dev.bwd = reinterpret_cast<const double *>(b + o_bwd);
            image.dev.tap_start = reinterpret_cast<const std::int32_t *>(b + o_ts);
            image.dev.blk_lo = reinterpret_cast<const std::int32_t *>(b + o_blk);
            image.ok = true;
        }
        image.built = true;
    }
    return image.ok ? &image.dev : nullptr;
}

const OperatorPlan *DevicePlan::operator_plan_locked() {
    if (!op_built_) {
        const trace::Range range(trace::Label::operator_build);
        auto op = std::make_unique<OperatorPlan>();
        if (build_operator_plan(*plan_, *op)) op_plan_ = std::move(op);
        op_built_ = true;
    }
    return op_plan_.get();
}

const OperatorPlan *DevicePlan::operator_plan() {
    std::lock_guard lock(mutex_);
    return operator_plan_locked();
}

int DevicePlan::built_op_steps() {
    std::unique_lock lock(mutex_, std::try_to_lock);
    if (!lock.owns_lock() || !op_built_) return -1;
    return op_plan_ ? op_plan_->steps : 0;
}

const DevOperator *DevicePlan::built_op(bool rows, bool bf16, int steps) {
    std::unique_lock lock(mutex_, std::try_to_lock);
    if (!lock.owns_lock() || !op_plan_ || !operator_fits(*op_plan_, rows, steps)) return nullptr;
    const auto it = op_images_.find((bf16 ? 8 : 0) + steps);
    return it != op_images_.end() && it->second.ok ? &it->second.dev : nullptr;
}

const DevOperator *DevicePlan::op(bool rows, bool bf16, int steps) {
    std::lock_guard lock(mutex_);
    const OperatorPlan *host = operator_plan_locked();
    if (!host || !operator_fits(*host, rows, steps)) return nullptr;
    OperatorImage &image = op_images_[(bf16 ? 8 : 0) + steps];
    if (!image.ok) {
        const std::vector<std::uint32_t> data = pack_operator(*host, bf16, steps);
        const std::size_t o_data = 0;
        const std::size_t o_k = align256(data.size() * 4);
        const std::size_t total = align256(o_k + host->kstart.size() * 4);
        std::vector<unsigned char> staging(total, 0);
        std::memcpy(staging.data() + o_data, data.data(), data.size() * 4);
        std::memcpy(staging.data() + o_k, host->kstart.data(), host->kstart.size() * 4);
        int current = 0;
        check(hipGetDevice(&current), "hipGetDevice failed");
        check(hipSetDevice(device_), "hipSetDevice failed");
        void *memory = nullptr;
        check(hipMalloc(&memory, total), "operator image allocation failed");
        if (const hipError_t e = upload(memory, staging.data(), total); e != hipSuccess) {
            (void)hipFree(memory);
            check(e, "operator image upload failed");
        }
        (void)hipSetDevice(current);
        grow(total);
        auto *b = static_cast<unsigned char *>(memory);
        image.memory = memory;
        image.dev.data = reinterpret_cast<const std::uint32_t *>(b + o_data);
        image.dev.kstart = reinterpret_cast<const std::int32_t *>(b + o_k);
        image.dev.n = host->n;
        image.dev.m = host->m;
        image.dev.blocks = host->blocks;
        image.dev.steps = steps;
        image.ok = true;
    }
    return &image.dev;
}

const DevRefine *DevicePlan::refine() {
    std::lock_guard lock(mutex_);
    if (!refine_built_) {
        const AxisPlan &p = *plan_;
        if (!p.requires_float64()) {
            refine_built_ = true;
            return nullptr;
        }
        const int n = p.destination_size, w = p.half_bandwidth, span = 2 * w + 1;
        const std::size_t nnz = p.transpose_weights_f64.size();
        const auto un = static_cast<std::size_t>(n);
        std::vector<float> w_hi(nnz), w_lo(nnz), g_hi(un * span, 0.0F), g_lo(un * span, 0.0F);
        auto split = [](double v, float &hi, float &lo) {
            hi = static_cast<float>(v);
            lo = static_cast<float>(v - static_cast<double>(hi));
        };
        for (std::size_t q = 0; q < nnz; ++q) split(p.transpose_weights_f64[q], w_hi[q], w_lo[q]);
        // normal_bands_f64[d * n + i] = G(i, i + d) (reference layout).
        for (int i = 0; i < n; ++i)
            for (int d = -w; d <= w; ++d) {
                const int j = i + d;
                if (j < 0 || j >= n) continue;
                const int lo_index = std::min(i, j);
                const double g = p.normal_bands_f64[static_cast<std::size_t>(std::abs(d)) * un
                                                    + static_cast<std::size_t>(lo_index)];
                const std::size_t at = static_cast<std::size_t>(i) * span + static_cast<std::size_t>(d + w);
                split(g, g_hi[at], g_lo[at]);
            }
        // Tap-major copies: k = the widest run of source samples a row spans.
        int k = 1;
        std::vector<std::int32_t> first(un, 0);
        for (int i = 0; i < n; ++i) {
            const auto k0 = p.transpose_offsets[static_cast<std::size_t>(i)];
            const auto k1 = p.transpose_offsets[static_cast<std::size_t>(i) + 1];
            if (k1 <= k0) continue;
            first[static_cast<std::size_t>(i)] = p.transpose_indices[k0];
            k = std::max(k, p.transpose_indices[k1 - 1] - p.transpose_indices[k0] + 1);
        }
        const auto uk = static_cast<std::size_t>(k);
        std::vector<double> t64(uk * un, 0.0), b64(un * span, 0.0);
        for (int i = 0; i < n; ++i) {
            const auto ui = static_cast<std::size_t>(i);
            for (auto q = p.transpose_offsets[ui]; q < p.transpose_offsets[ui + 1]; ++q) {
                const auto e = static_cast<std::size_t>(p.transpose_indices[q] - first[ui]);
                t64[e * un + ui] = p.transpose_weights_f64[q];
            }
            for (int d = -w; d <= w; ++d) {
                const int j = i + d;
                if (j < 0 || j >= n) continue;
                b64[static_cast<std::size_t>(d + w) * un + ui]
                    = p.normal_bands_f64[static_cast<std::size_t>(std::abs(d)) * un
                                         + static_cast<std::size_t>(std::min(i, j))];
            }
        }
        const std::size_t o_wh = 0;
        const std::size_t o_wl = align256(o_wh + std::max<std::size_t>(nnz, 1) * 4);
        const std::size_t o_gh = align256(o_wl + std::max<std::size_t>(nnz, 1) * 4);
        const std::size_t o_gl = align256(o_gh + g_hi.size() * 4);
        const std::size_t o_fi = align256(o_gl + g_lo.size() * 4);
        const std::size_t o_t = align256(o_fi + first.size() * 4);
        const std::size_t o_b = align256(o_t + t64.size() * 8);
        const std::size_t total = align256(o_b + b64.size() * 8);
        std::vector<unsigned char> host(total, 0);
        std::memcpy(host.data() + o_wh, w_hi.data(), nnz * 4);
        std::memcpy(host.data() + o_wl, w_lo.data(), nnz * 4);
        std::memcpy(host.data() + o_gh, g_hi.data(), g_hi.size() * 4);
        std::memcpy(host.data() + o_gl, g_lo.data(), g_lo.size() * 4);
        std::memcpy(host.data() + o_fi, first.data(), first.size() * 4);
        std::memcpy(host.data() + o_t, t64.data(), t64.size() * 8);
        std::memcpy(host.data() + o_b, b64.data(), b64.size() * 8);
        int current = 0;
        check(hipGetDevice(&current), "hipGetDevice failed");
        check(hipSetDevice(device_), "hipSetDevice failed");
        void *memory = nullptr;
        check(hipMalloc(&memory, total), "refinement table allocation failed");
        if (const hipError_t e = upload(memory, host.data(), total);
            e != hipSuccess) {
            (void)hipFree(memory);
            check(e, "refinement table upload failed");
        }
        (void)hipSetDevice(current);
        refine_memory_ = memory;
        refine_bytes_ = total;
        grow(total);
        refine_built_ = true;
        auto *b = static_cast<unsigned char *>(refine_memory_);
        refine_.n = n;
        refine_.w = w;
        refine_.w_hi = reinterpret_cast<const float *>(b + o_wh);
        refine_.w_lo = reinterpret_cast<const float *>(b + o_wl);
        refine_.g_hi = reinterpret_cast<const float *>(b + o_gh);
        refine_.g_lo = reinterpret_cast<const float *>(b + o_gl);
        refine_.m = p.source_size;
        refine_.k = k;
        refine_.first = reinterpret_cast<const std::int32_t *>(b + o_fi);
        refine_.t64 = reinterpret_cast<const double *>(b + o_t);
        refine_.b64 = reinterpret_cast<const double *>(b + o_b);
    }
    return refine_memory_ ? &refine_ : nullptr;
}

} // namespace dsmvc::hip
