// Packing of axis plans into device images and chunk partitions.
#include "device_plan.hpp"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

#include "kernels.hpp"

namespace dsmvc::hip {
namespace {

void check(hipError_t e, const char *what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string("dsmvc: ") + what + ": " + hipGetErrorString(e));
}

std::size_t align256(std::size_t v) { return (v + 255U) & ~std::size_t{255}; }

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

bool build_chunk_plan(const AxisPlan &p, int chunks, ChunkPlan &out) {
    const int n = p.destination_size, w = p.half_bandwidth;
    if (!chunked_band_supported(w) || chunks < 1) return false;
    const int length = (n + chunks - 1) / chunks;
    if (length < w || length > chunked_max_length()) return false;
    const int count = (n + length - 1) / length;
    out.chunks = count;
    out.length = length;
    out.tf.assign(static_cast<std::size_t>(count) * w * w, 0.0F);
    out.tb.assign(static_cast<std::size_t>(count) * w * w, 0.0F);
    out.win.assign(static_cast<std::size_t>(count) * 2, 0);
    const auto un = static_cast<std::size_t>(n);
    auto lld = [&](int d, int i) { return static_cast<double>(p.lower_ld[static_cast<std::size_t>(d - 1) * un + static_cast<std::size_t>(i)]); };
    auto ul = [&](int d, int i) { return static_cast<double>(p.upper_l[static_cast<std::size_t>(d - 1) * un + static_cast<std::size_t>(i)]); };
    std::vector<double> ring(static_cast<std::size_t>(w));
    auto push = [&](double v) {
        for (int d = w - 1; d > 0; --d) ring[static_cast<std::size_t>(d)] = ring[static_cast<std::size_t>(d - 1)];
        ring[0] = v;
    };
    out.ring_feasible = true;
    const int span = rows_ring_span();
    for (int q = 0; q < count; ++q) {
        const int a = q * length, b = std::min(n, a + length);
        for (int c = 0; c < w; ++c) {
            std::fill(ring.begin(), ring.end(), 0.0);
            ring[static_cast<std::size_t>(c)] = 1.0;
            for (int i = a; i < b; ++i) {
                double v = 0.0;
                for (int d = w; d >= 1; --d) v -= lld(d, i) * ring[static_cast<std::size_t>(d - 1)];
                push(v * static_cast<double>(p.inverse_diagonal[static_cast<std::size_t>(i)]));
            }
            for (int r = 0; r < w; ++r)
                out.tf[(static_cast<std::size_t>(q) * w + r) * w + c] = static_cast<float>(ring[static_cast<std::size_t>(r)]);
            std::fill(ring.begin(), ring.end(), 0.0);
            ring[static_cast<std::size_t>(c)] = 1.0;
            for (int i = b - 1; i >= a; --i) {
                double v = 0.0;
                for (int d = w; d >= 1; --d) v -= ul(d, i) * ring[static_cast<std::size_t>(d - 1)];
                push(v);
            }
            for (int r = 0; r < w; ++r)
                out.tb[(static_cast<std::size_t>(q) * w + r) * w + c] = static_cast<float>(ring[static_cast<std::size_t>(r)]);
        }
        // Source window and the rows kernel's staging walk (a ring of two
        // 16-column halves that only moves forward).
        int lo = p.source_size, hi = 0;
        for (int i = a; i < b; ++i) {
            const auto k0 = p.transpose_offsets[static_cast<std::size_t>(i)];
            const auto k1 = p.transpose_offsets[static_cast<std::size_t>(i) + 1];
            if (k1 > k0) {
                lo = std::min(lo, p.transpose_indices[k0]);
                hi = std::max(hi, p.transpose_indices[k1 - 1]);
            }
        }
        if (lo > hi) lo = hi = 0;
        out.win[static_cast<std::size_t>(q) * 2] = lo;
        out.win[static_cast<std::size_t>(q) * 2 + 1] = hi;
        int base = lo & ~(span - 1);
        for (int i = a; i < b; ++i) {
            const auto k0 = p.transpose_offsets[static_cast<std::size_t>(i)];
            const auto k1 = p.transpose_offsets[static_cast<std::size_t>(i) + 1];
            if (k1 == k0) continue;
            const int top = p.transpose_indices[k1 - 1];
            while (top >= base + 2 * span) base += span;
            if (p.transpose_indices[k0] < base) out.ring_feasible = false;
        }
    }
    return true;
}

int choose_chunks(const AxisPlan &p, bool rows) {
    const int n = p.destination_size, w = p.half_bandwidth;
    if (!chunked_band_supported(w)) return 0;
    for (const int c : {16, 12, 8}) {
        const int length = (n + c - 1) / c;
        if (length < std::max(w, 8) || length > chunked_max_length()) continue;
        (void)rows;
        return c;
    }
    return 0;
}

DevicePlan::DevicePlan(const AxisPlan &plan, int device)
    : keep_(std::make_shared<const AxisPlan>(plan)), plan_(keep_.get()), device_(device) {
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
    const hipError_t copy = hipMemcpy(memory_, host.data(), bytes_, hipMemcpyHostToDevice);
    if (copy != hipSuccess) {
        (void)hipFree(memory_);
        memory_ = nullptr;
        check(copy, "device plan upload failed");
    }
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

DevicePlan::~DevicePlan() {
    for (auto &entry : chunk_images_)
        if (entry.second.memory) (void)hipFree(entry.second.memory);
    if (memory_) (void)hipFree(memory_);
}

const DevChunks *DevicePlan::chunks(int count, bool rows) {
    std::lock_guard lock(mutex_);
    auto it = chunk_images_.find(count);
    if (it == chunk_images_.end()) {
        ChunkImage image;
        ChunkPlan cp;
        if (build_chunk_plan(*plan_, count, cp)) {
            const std::size_t mats = cp.tf.size() * 4;
            const std::size_t o_tb = align256(mats);
            const std::size_t o_win = align256(o_tb + mats);
            const std::size_t total = align256(o_win + cp.win.size() * 4);
            std::vector<unsigned char> host(total, 0);
            std::memcpy(host.data(), cp.tf.data(), mats);
            std::memcpy(host.data() + o_tb, cp.tb.data(), mats);
            std::memcpy(host.data() + o_win, cp.win.data(), cp.win.size() * 4);
            int current = 0;
            check(hipGetDevice(&current), "hipGetDevice failed");
            check(hipSetDevice(device_), "hipSetDevice failed");
            check(hipMalloc(&image.memory, total), "chunk table allocation failed");
            check(hipMemcpy(image.memory, host.data(), total, hipMemcpyHostToDevice),
                  "chunk table upload failed");
            (void)hipSetDevice(current);
            auto *b = static_cast<unsigned char *>(image.memory);
            image.dev.chunks = cp.chunks;
            image.dev.length = cp.length;
            image.dev.tf = reinterpret_cast<const float *>(b);
            image.dev.tb = reinterpret_cast<const float *>(b + o_tb);
            image.dev.win = reinterpret_cast<const std::int32_t *>(b + o_win);
            image.ok = true;
            image.ring_feasible = cp.ring_feasible;
        }
        it = chunk_images_.emplace(count, image).first;
    }
    if (!it->second.ok || (rows && !it->second.ring_feasible)) return nullptr;
    return &it->second.dev;
}

} // namespace dsmvc::hip
