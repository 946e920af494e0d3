// FP64 host planner for the HIP descale executor.
//
// Produces the same AxisPlan as the reference planner
// (reference src/axis_plan.cpp:253-970), bit for bit: every floating-point
// expression below keeps the reference's operands and evaluation order,
// because the F32 factors and the CSR weights are compared exactly.  The
// organisation is this library's own: one pass per source sample builds the
// coalesced operator row, a counting transpose gives the destination-major
// CSR, and a BandSystem object owns formation, factorisation and the
// conditioning test.  Compile with -ffp-contract=off.
#include "engine.hpp"
#include "trace.hpp"

#include <algorithm>
#include <atomic>
#include <bit>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <list>
#include <mutex>
#include <numbers>
#include <unordered_map>

namespace dsmvc {
namespace {

constexpr double kEps = std::numeric_limits<double>::epsilon();
constexpr double kF64Threshold = 1.0e-4;   // rcond below this needs FP64 (axis_plan.cpp:893)

// ---------------------------------------------------------------------------
// Filter kernels.  Polynomials and their evaluation order are those of the
// reference's filter_weight (axis_plan.cpp:272-344); they must match exactly.
// ---------------------------------------------------------------------------
double sq(double v) { return v * v; }
double cb(double v) { return v * v * v; }

double sinc_pi(double v) {
    if (v == 0.0) return 1.0;
    const double a = std::numbers::pi * v;
    return std::sin(a) / a;
}

double kernel_weight(const KernelSpec &k, double distance) {
    double x = std::abs(distance);
    switch (k.kind) {
    case KernelKind::bilinear:
        return std::max(1.0 - x, 0.0);
    case KernelKind::bicubic:
        if (x < 1.0)
            return ((12.0 - 9.0 * k.b - 6.0 * k.c) * cb(x)
                    + (-18.0 + 12.0 * k.b + 6.0 * k.c) * sq(x) + (6.0 - 2.0 * k.b)) / 6.0;
        if (x < 2.0)
            return ((-k.b - 6.0 * k.c) * cb(x) + (6.0 * k.b + 30.0 * k.c) * sq(x)
                    + (-12.0 * k.b - 48.0 * k.c) * x + (8.0 * k.b + 24.0 * k.c)) / 6.0;
        return 0.0;
    case KernelKind::lanczos:
        return k.taps > 0 && x < static_cast<double>(k.taps)
            ? sinc_pi(x) * sinc_pi(x / static_cast<double>(k.taps)) : 0.0;
    case KernelKind::spline16:
        if (x < 1.0) return 1.0 - x / 5.0 - 9.0 * sq(x) / 5.0 + cb(x);
        if (x < 2.0) {
            x -= 1.0;
            return -7.0 * x / 15.0 + 4.0 * sq(x) / 5.0 - cb(x) / 3.0;
        }
        return 0.0;
    case KernelKind::spline36:
        if (x < 1.0) return 1.0 - 3.0 * x / 209.0 - 453.0 * sq(x) / 209.0 + 13.0 * cb(x) / 11.0;
        if (x < 2.0) {
            x -= 1.0;
            return -156.0 * x / 209.0 + 270.0 * sq(x) / 209.0 - 6.0 * cb(x) / 11.0;
        }
        if (x < 3.0) {
            x -= 2.0;
            return 26.0 * x / 209.0 - 45.0 * sq(x) / 209.0 + cb(x) / 11.0;
        }
        return 0.0;
    case KernelKind::spline64:
        if (x < 1.0)
            return 1.0 - 3.0 * x / 2911.0 - 6387.0 * sq(x) / 2911.0 + 49.0 * cb(x) / 41.0;
        if (x < 2.0) {
            x -= 1.0;
            return -2328.0 * x / 2911.0 + 4032.0 * sq(x) / 2911.0 - 24.0 * cb(x) / 41.0;
        }
        if (x < 3.0) {
            x -= 2.0;
            return 582.0 * x / 2911.0 - 1008.0 * sq(x) / 2911.0 + 6.0 * cb(x) / 41.0;
        }
        if (x < 4.0) {
            x -= 3.0;
            return -97.0 * x / 2911.0 + 168.0 * sq(x) / 2911.0 - cb(x) / 41.0;
        }
        return 0.0;
    case KernelKind::custom:
        return 0.0;
    }
    return 0.0;
}

// Validation and effective support (reference filter_support, axis_plan.cpp:370-424).
std::int32_t effective_support(const AxisRequest &r) {
    if (r.source_size <= 0 || r.destination_size <= 0)
        throw std::invalid_argument("axis dimensions must be positive");
    if (!(r.active_length > 0.0) || !std::isfinite(r.active_length) || !std::isfinite(r.shift))
        throw std::invalid_argument(
            "active length and shift must be finite, with positive active length");
    if (static_cast<unsigned>(r.border) > static_cast<unsigned>(BorderMode::mirror))
        throw std::invalid_argument("padding mode is invalid");
    if (static_cast<unsigned>(r.f64_mode) > static_cast<unsigned>(F64Mode::float64_only))
        throw std::invalid_argument("f64 mode is invalid");
    if (!(r.kernel.blur > 0.0) || !std::isfinite(r.kernel.blur))
        throw std::invalid_argument("blur must be finite and greater than zero");
    std::int32_t base = 0;
    switch (r.kernel.kind) {
    case KernelKind::bilinear: base = 1; break;
    case KernelKind::bicubic:
    case KernelKind::spline16: base = 2; break;
    case KernelKind::spline36: base = 3; break;
    case KernelKind::spline64: base = 4; break;
    case KernelKind::lanczos:
    case KernelKind::custom: base = r.kernel.taps; break;
    default: throw std::invalid_argument("kernel is invalid");
    }
    constexpr auto max_support = (std::numeric_limits<std::int32_t>::max() - 1) / 2;
    if (base <= 0 || base > max_support)
        throw std::invalid_argument("filter support is invalid or too large");
    const double scaled = static_cast<double>(base) * r.kernel.blur;
    if (!std::isfinite(scaled) || scaled > static_cast<double>(max_support))
        throw std::length_error("effective filter support is too large");
    const double effective = std::ceil(scaled);
    if (!(effective >= 1.0) || effective > static_cast<double>(max_support))
        throw std::length_error("effective filter support is too large");
    return static_cast<std::int32_t>(effective);
}

// Sparse matrix in compressed rows with FP64 values.
struct SparseRows {
    std::vector<std::uint32_t> start;
    std::vector<std::int32_t> column;
    std::vector<double> value;
};

// Sampling geometry of an axis: for every source sample, its 2*support tap
// offsets (tap centre - sample position) and the destination index of each
// tap through the border rule (-1: none), plus the sample's distinct
// destination indices in ascending order.  It depends only on the sizes,
// active length, shift, support and border, so every kernel and blur of the
// same support reuses it (reference build_geometry, axis_plan.cpp:456-518,
// cached by get_geometry, 520-533).
struct Geometry {
    std::int32_t taps = 0;
    std::vector<double> offset;          // [sample][tap]
    std::vector<std::int32_t> dest;      // [sample][tap]
    std::vector<std::uint32_t> start;    // [sample + 1] into `unique`
    std::vector<std::int32_t> unique;

    [[nodiscard]] std::size_t storage_bytes() const noexcept {
        return sizeof(*this) + offset.capacity() * sizeof(double) + dest.capacity() * 4
            + start.capacity() * 4 + unique.capacity() * 4;
    }
};

Geometry build_geometry(const AxisRequest &r, std::int32_t support) {
    const trace::Range range(trace::Label::geometry_build);
    const std::int32_t taps = 2 * support;
    const auto m = static_cast<std::size_t>(r.source_size);
    if (m > std::numeric_limits<std::size_t>::max() / 2 / static_cast<std::size_t>(taps))
        throw std::length_error("inverse geometry is too large");
    Geometry g;
    g.taps = taps;
    g.offset.resize(m * static_cast<std::size_t>(taps));
    g.dest.resize(m * static_cast<std::size_t>(taps));
    g.start.reserve(m + 1);
    g.unique.reserve(m * static_cast<std::size_t>(taps));
    g.start.push_back(0U);
    std::vector<std::int32_t> dests;
    dests.reserve(static_cast<std::size_t>(taps));
    const double ratio = static_cast<double>(r.source_size) / r.active_length;
    for (std::int32_t s = 0; s < r.source_size; ++s) {
        const double position = (static_cast<double>(s) + 0.5) / ratio + r.shift;
        const double lifted = position - static_cast<double>(support);
        const double first = (lifted < 0.0 ? std::floor(lifted + 0.5)
                                           : std::floor(lifted + 0.49999999999999994)) + 0.5;
        const std::size_t base = static_cast<std::size_t>(s) * static_cast<std::size_t>(taps);
        dests.clear();
        for (std::int32_t t = 0; t < taps; ++t) {
            const double center = first + static_cast<double>(t);
            g.offset[base + static_cast<std::size_t>(t)] = center - position;
            const double cell = std::floor(center);
            if (!std::isfinite(cell) || cell < -0x1p63 || cell >= 0x1p63)
                throw std::out_of_range(
                    "shift places filter support outside the 64-bit pixel grid");
            const std::int32_t d =
                padding_index(static_cast<std::int64_t>(cell), r.destination_size, r.border);
            g.dest[base + static_cast<std::size_t>(t)] = d;
            if (d >= 0 && std::find(dests.begin(), dests.end(), d) == dests.end()) dests.push_back(d);
        }
        std::sort(dests.begin(), dests.end());
        g.unique.insert(g.unique.end(), dests.begin(), dests.end());
        if (g.unique.size() > std::numeric_limits<std::uint32_t>::max())
            throw std::length_error("inverse geometry coefficient table is too large");
        g.start.push_back(static_cast<std::uint32_t>(g.unique.size()));
    }
    return g;
}

// Bounded LRU of geometries keyed by the exact bits of what they depend on
// (reference geometry_key, axis_plan.cpp:444-454).  A geometry is built
// outside the lock; two threads that miss the same key at once both build
// it and the first insertion wins (the results are identical).
class GeometryCache {
public:
    std::shared_ptr<const Geometry> get(const AxisRequest &r, std::int32_t support) {
        const Key key{static_cast<std::uint32_t>(r.source_size),
                      static_cast<std::uint32_t>(r.destination_size),
                      std::bit_cast<std::uint64_t>(r.active_length),
                      std::bit_cast<std::uint64_t>(r.shift), static_cast<std::uint32_t>(support),
                      static_cast<std::uint32_t>(r.border)};
        {
            std::lock_guard lock(mutex_);
            if (auto it = map_.find(key); it != map_.end()) {
                ++hits_;
                lru_.splice(lru_.begin(), lru_, it->second.where);
                return it->second.value;
            }
        }
        auto built = std::make_shared<const Geometry>(build_geometry(r, support));
        std::lock_guard lock(mutex_);
        ++builds_;
        if (auto it = map_.find(key); it != map_.end()) return it->second.value;
        lru_.push_front(key);
        bytes_ += built->storage_bytes();
        map_.emplace(key, Entry{built, lru_.begin()});
        while ((map_.size() > kEntries || bytes_ > kBytes) && map_.size() > 1) {
            auto victim = map_.find(lru_.back());
            bytes_ -= victim->second.value->storage_bytes();
            map_.erase(victim);
            lru_.pop_back();
        }
        return built;
    }

    void stats(PlannerCacheStats &s) {
        std::lock_guard lock(mutex_);
        s.geometry_hits = hits_;
        s.geometry_builds = builds_;
        s.geometry_entries = map_.size();
        s.geometry_resident_bytes = bytes_;
    }

    void clear() {
        std::lock_guard lock(mutex_);
        map_.clear();
        lru_.clear();
        bytes_ = 0;
        hits_ = builds_ = 0;
    }

private:
    struct Key {
        std::uint32_t source, destination;
        std::uint64_t active, shift;
        std::uint32_t support, border;
        bool operator==(const Key &) const = default;
    };
    struct KeyHash {
        std::size_t operator()(const Key &k) const noexcept {
            std::uint64_t h = 0x9E3779B97F4A7C15ULL;
            for (const std::uint64_t w : {std::uint64_t{k.source} << 32 | k.destination, k.active,
                                          k.shift, std::uint64_t{k.support} << 32 | k.border})
                h = (h ^ w) * 0xBF58476D1CE4E5B9ULL + (h >> 31);
            return static_cast<std::size_t>(h);
        }
    };
    struct Entry {
        std::shared_ptr<const Geometry> value;
        std::list<Key>::iterator where;
    };
    static constexpr std::size_t kEntries = 1024;
    static constexpr std::size_t kBytes = 128U << 20;
    std::mutex mutex_;
    std::unordered_map<Key, Entry, KeyHash> map_;
    std::list<Key> lru_;
    std::size_t bytes_ = 0;
    std::uint64_t hits_ = 0, builds_ = 0;
};

GeometryCache &geometry_cache() {
    static GeometryCache cache;
    return cache;
}

// The forward resampling operator A (source samples x destination samples)
// on a geometry: one row per source sample, its taps' kernel weights
// normalised by their sum and coalesced per destination index in ascending
// index order (reference make_descale_matrix, axis_plan.cpp:535-599).
SparseRows sample_operator(const AxisRequest &r, const Geometry &g, const CustomKernel &custom) {
    const std::int32_t taps = g.taps;
    const auto m = static_cast<std::size_t>(r.source_size);
    SparseRows a;
    a.start.reserve(m + 1);
    a.column.reserve(g.unique.size());
    a.value.reserve(g.unique.size());
    a.start.push_back(0U);

    std::vector<double> weight(static_cast<std::size_t>(taps));
    std::vector<double> merged;
    std::vector<char> used;
    const bool unit_blur = r.kernel.blur == 1.0;
    for (std::size_t s = 0; s < m; ++s) {
        const std::size_t base = s * static_cast<std::size_t>(taps);
        double total = 0.0;
        for (std::int32_t t = 0; t < taps; ++t) {
            const double offset = g.offset[base + static_cast<std::size_t>(t)];
            const double d = unit_blur ? offset : offset / r.kernel.blur;
            // A custom kernel is evaluated at |d| (reference make_descale_matrix,
            // axis_plan.cpp:560-563).
            const double w = r.kernel.kind == KernelKind::custom ? custom(std::abs(d))
                                                                 : kernel_weight(r.kernel, d);
            if (!std::isfinite(w)) throw std::runtime_error("filter returned a non-finite value");
            weight[static_cast<std::size_t>(t)] = w;
            total += w;
        }
        if (!std::isfinite(total) || total == 0.0)
            throw std::runtime_error("filter produced a zero or non-finite weight sum");
        const auto ub = g.unique.begin() + g.start[s], ue = g.unique.begin() + g.start[s + 1];
        const auto count = static_cast<std::size_t>(ue - ub);
        merged.assign(count, 0.0);
        used.assign(count, 0);
        for (std::int32_t t = 0; t < taps; ++t) {
            const std::int32_t dest = g.dest[base + static_cast<std::size_t>(t)];
            if (dest < 0) continue;
            const auto slot = static_cast<std::size_t>(std::lower_bound(ub, ue, dest) - ub);
            const double w = weight[static_cast<std::size_t>(t)] / total;
            if (w == 0.0) continue;
            merged[slot] += w;
            used[slot] = 1;
        }
        for (std::size_t k = 0; k < count; ++k) {
            if (!used[k]) continue;
            a.column.push_back(ub[static_cast<std::ptrdiff_t>(k)]);
            a.value.push_back(merged[k]);
        }
        a.start.push_back(static_cast<std::uint32_t>(a.column.size()));
    }
    return a;
}

// Counting transpose: rows of A^T are destination samples, their source
// indices ascending because source rows are visited in order.
SparseRows transpose(const SparseRows &a, std::int32_t rows, std::int32_t columns) {
    SparseRows t;
    t.start.assign(static_cast<std::size_t>(columns) + 1, 0U);
    for (const auto c : a.column) ++t.start[static_cast<std::size_t>(c) + 1];
    for (std::size_t c = 1; c < t.start.size(); ++c) t.start[c] += t.start[c - 1];
    t.column.resize(a.column.size());
    t.value.resize(a.value.size());
    std::vector<std::uint32_t> fill(t.start.begin(), t.start.end() - 1);
    for (std::int32_t s = 0; s < rows; ++s) {
        for (auto k = a.start[static_cast<std::size_t>(s)]; k < a.start[static_cast<std::size_t>(s) + 1]; ++k) {
            const auto at = fill[static_cast<std::size_t>(a.column[k])]++;
            t.column[at] = s;
            t.value[at] = a.value[k];
        }
    }
    return t;
}

// Symmetric banded normal system G = A^T A, stored by diagonal distance:
// band[d * n + i] = G(i, i + d).  Factorised in place as G = L D L^T with
// the pivot regularised by +DBL_EPSILON (axis_plan.cpp:766-791).
class BandSystem {
public:
    BandSystem(std::int32_t n, std::int32_t w)
        : n_(n), w_(w), band_((static_cast<std::size_t>(w) + 1) * static_cast<std::size_t>(n), 0.0) {}

    [[nodiscard]] double &at(std::int32_t d, std::int32_t i) {
        return band_[static_cast<std::size_t>(d) * static_cast<std::size_t>(n_) + static_cast<std::size_t>(i)];
    }
    [[nodiscard]] double at(std::int32_t d, std::int32_t i) const {
        return band_[static_cast<std::size_t>(d) * static_cast<std::size_t>(n_) + static_cast<std::size_t>(i)];
    }
    [[nodiscard]] const std::vector<double> &data() const { return band_; }
    [[nodiscard]] std::vector<double> &data() { return band_; }

    // Dot products of A^T rows over shared source indices, ascending
    // (axis_plan.cpp:628-668).
    void form(const SparseRows &t) {
        for (std::int32_t i = 0; i < n_; ++i) {
            const std::int32_t reach = std::min(w_, n_ - i - 1);
            for (std::int32_t d = 0; d <= reach; ++d) {
                auto p = t.start[static_cast<std::size_t>(i)];
                const auto pe = t.start[static_cast<std::size_t>(i) + 1];
                auto q = t.start[static_cast<std::size_t>(i + d)];
                const auto qe = t.start[static_cast<std::size_t>(i + d) + 1];
                double dot = 0.0;
                while (p < pe && q < qe) {
                    const auto sp = t.column[p], sq_ = t.column[q];
                    if (sp < sq_) ++p;
                    else if (sq_ < sp) ++q;
                    else dot += t.value[p++] * t.value[q++];
                }
                at(d, i) = dot;
            }
        }
    }

    void factor() {
        for (std::int32_t i = 0; i < n_; ++i) {
            const std::int32_t end = std::min(w_ + 1, n_ - i);
            const double pivot = at(0, i) + kEps;
            for (std::int32_t d = 1; d < end; ++d) {
                const double mult = at(d, i) / pivot;
                for (std::int32_t o = 0; o < end - d; ++o) at(o, i + d) -= mult * at(d + o, i);
            }
            const double inv = 1.0 / pivot;
            for (std::int32_t d = 1; d < end; ++d) at(d, i) *= inv;
        }
    }

    // Unit-lower / diagonal / unit-upper substitution with the factors
    // (axis_plan.cpp:793-823); used by the condition estimate.
    void solve(std::vector<double> &v) const {
        for (std::int32_t i = 0; i < n_; ++i) {
            double s = v[static_cast<std::size_t>(i)];
            for (std::int32_t d = std::min(w_, i); d >= 1; --d) s -= at(d, i - d) * v[static_cast<std::size_t>(i - d)];
            v[static_cast<std::size_t>(i)] = s;
        }
        for (std::int32_t i = 0; i < n_; ++i) v[static_cast<std::size_t>(i)] /= at(0, i) + kEps;
        for (std::int32_t i = n_ - 2; i >= 0; --i) {
            double s = v[static_cast<std::size_t>(i)];
            for (std::int32_t d = std::min(w_, n_ - i - 1); d >= 1; --d) s -= at(d, i) * v[static_cast<std::size_t>(i + d)];
            v[static_cast<std::size_t>(i)] = s;
        }
    }

private:
    std::int32_t n_;
    std::int32_t w_;
    std::vector<double> band_;
};

double unit_clamp(double v) { return std::clamp(v, 0.0, 1.0); }

// Conditioning of G: a diagonal-dominance bound first, then the comparison-
// matrix bound on the factors, then an 8-step Hager 1-norm estimate; each
// runs only when the previous one cannot clear the FP64 threshold
// (axis_plan.cpp:675-873, 893-908).
struct Conditioning {
    double one_norm = 0.0;
    double rcond = 0.0;
};

Conditioning condition(const BandSystem &g, const BandSystem &f, std::int32_t n, std::int32_t w) {
    std::vector<double> acc(static_cast<std::size_t>(n), 0.0);
    for (std::int32_t c = 0; c < n; ++c) {
        acc[static_cast<std::size_t>(c)] += std::abs(g.at(0, c));
        for (std::int32_t d = 1, reach = std::min(w, n - c - 1); d <= reach; ++d) {
            const double v = std::abs(g.at(d, c));
            acc[static_cast<std::size_t>(c)] += v;
            acc[static_cast<std::size_t>(c + d)] += v;
        }
    }
    Conditioning out;
    out.one_norm = *std::max_element(acc.begin(), acc.end());
    double margin = std::numeric_limits<double>::infinity();
    for (std::int32_t i = 0; i < n; ++i)
        margin = std::min(margin, 2.0 * std::abs(g.at(0, i)) - acc[static_cast<std::size_t>(i)]);
    out.rcond = margin > 0.0 && out.one_norm > 0.0 ? margin / out.one_norm : 0.0;
    if (out.rcond >= kF64Threshold) return out;

    // |L^-1| <= (I - |E|)^-1 row and column sums.
    std::fill(acc.begin(), acc.end(), 1.0);
    double linf = 1.0, dinv = 0.0;
    for (std::int32_t i = 0; i < n; ++i) {
        double s = 1.0;
        for (std::int32_t d = 1, reach = std::min(w, i); d <= reach; ++d)
            s += std::abs(f.at(d, i - d)) * acc[static_cast<std::size_t>(i - d)];
        acc[static_cast<std::size_t>(i)] = s;
        linf = std::max(linf, s);
        dinv = std::max(dinv, 1.0 / std::abs(f.at(0, i) + kEps));
    }
    std::fill(acc.begin(), acc.end(), 1.0);
    double l1 = 1.0;
    for (std::int32_t c = n - 1; c >= 0; --c) {
        double s = 1.0;
        for (std::int32_t d = 1, reach = std::min(w, n - c - 1); d <= reach; ++d)
            s += std::abs(f.at(d, c)) * acc[static_cast<std::size_t>(c + d)];
        acc[static_cast<std::size_t>(c)] = s;
        l1 = std::max(l1, s);
    }
    const double bound = l1 * dinv * linf;
    out.rcond = (!(out.one_norm > 0.0) || !(bound > 0.0) || !std::isfinite(out.one_norm)
                 || !std::isfinite(bound)) ? 0.0 : unit_clamp(1.0 / (out.one_norm * bound));
    if (out.rcond >= kF64Threshold) return out;

    // Hager's estimate of ||G^-1||_1.
    const auto size = static_cast<std::size_t>(n);
    std::vector<double> x(size, 1.0 / static_cast<double>(n)), y(size), z(size);
    double estimate = 0.0;
    std::size_t previous = size;
    for (int step = 0; step < 8; ++step) {
        y = x;
        f.solve(y);
        double norm = 0.0;
        for (const double v : y) norm = norm + std::abs(v);
        estimate = std::max(estimate, norm);
        for (std::size_t i = 0; i < size; ++i) z[i] = y[i] < 0.0 ? -1.0 : 1.0;
        f.solve(z);
        std::size_t arg = 0;
        for (std::size_t i = 1; i < size; ++i)
            if (std::abs(z[arg]) < std::abs(z[i])) arg = i;
        double dot = 0.0;
        for (std::size_t i = 0; i < size; ++i) dot = dot + z[i] * x[i];
        if (arg == previous || std::abs(z[arg]) <= dot) break;
        std::fill(x.begin(), x.end(), 0.0);
        x[arg] = 1.0;
        previous = arg;
    }
    out.rcond = (!(out.one_norm > 0.0) || !(estimate > 0.0) || !std::isfinite(out.one_norm)
                 || !std::isfinite(estimate)) ? 0.0 : unit_clamp(1.0 / (out.one_norm * estimate));
    return out;
}

template <class T>
bool finite_all(const std::vector<T> &v) {
    return std::all_of(v.begin(), v.end(), [](T x) { return std::isfinite(x); });
}

} // namespace

std::int32_t padding_index(std::int64_t index, std::int32_t size, BorderMode mode) {
    if (size <= 0) throw std::invalid_argument("padding size must be positive");
    if (static_cast<unsigned>(mode) > static_cast<unsigned>(BorderMode::mirror))
        throw std::invalid_argument("padding mode is invalid");
    const std::int64_t n = size;
    if (index >= 0 && index < n) return static_cast<std::int32_t>(index);
    switch (mode) {
    case BorderMode::zero:
        return -1;
    case BorderMode::repeat:
        return index < 0 ? 0 : size - 1;
    case BorderMode::reflect101: {
        if (size == 1) return 0;
        const std::int64_t period = 2 * (n - 1);
        std::int64_t k = index % period;
        if (k < 0) k += period;
        return static_cast<std::int32_t>(k < n ? k : period - k);
    }
    case BorderMode::symmetric: {
        const std::int64_t period = 2 * n;
        std::int64_t k = index % period;
        if (k < 0) k += period;
        return static_cast<std::int32_t>(k < n ? k : period - 1 - k);
    }
    case BorderMode::mirror:
        if (index < -n || index >= 2 * n) return -1;
        return static_cast<std::int32_t>(index < 0 ? -index - 1 : 2 * n - index - 1);
    }
    throw std::invalid_argument("padding mode is invalid");
}

bool AxisPlan::valid() const noexcept {
    const std::uint8_t known = checked.state.load(std::memory_order_acquire);
    if (known != 0) return known == 1;
    const bool ok = check_arrays();
    checked.state.store(ok ? 1 : 2, std::memory_order_release);
    return ok;
}

bool AxisPlan::check_arrays() const noexcept {
    if (source_size <= 0 || destination_size <= 0 || support <= 0 || half_bandwidth < 0
        || half_bandwidth >= destination_size || !(active_length > 0.0)
        || !std::isfinite(active_length) || !std::isfinite(shift) || !std::isfinite(normal_rcond)
        || normal_rcond < 0.0 || normal_rcond > 1.0 || !std::isfinite(normal_inf_norm)
        || normal_inf_norm < 0.0)
        return false;
    const auto n = static_cast<std::size_t>(destination_size);
    const auto factors = static_cast<std::size_t>(half_bandwidth) * n;
    if (transpose_offsets.size() != n + 1 || transpose_indices.size() != transpose_weights.size()
        || transpose_offsets.front() != 0U || transpose_offsets.back() != transpose_indices.size()
        || lower_ld.size() != factors || upper_l.size() != factors || inverse_diagonal.size() != n)
        return false;
    const bool f64 = !ldlt_bands_f64.empty();
    if (f64 != !transpose_weights_f64.empty() || f64 != !normal_bands_f64.empty()
        || f64 != !inverse_diagonal_f64.empty())
        return false;
    if (f64 ? !(normal_inf_norm > 0.0) : normal_inf_norm != 0.0) return false;
    const auto bands = (static_cast<std::size_t>(half_bandwidth) + 1) * n;
    if (f64 && (transpose_weights_f64.size() != transpose_weights.size()
                || normal_bands_f64.size() != bands || ldlt_bands_f64.size() != bands
                || inverse_diagonal_f64.size() != n))
        return false;
    for (std::size_t i = 0; i < n; ++i) {
        const auto b = transpose_offsets[i], e = transpose_offsets[i + 1];
        if (b > e || e > transpose_indices.size()) return false;
        std::int32_t prev = -1;
        for (auto k = b; k < e; ++k) {
            const auto s = transpose_indices[k];
            if (s < 0 || s >= source_size || s <= prev || !std::isfinite(transpose_weights[k]))
                return false;
            prev = s;
        }
    }
    if (!finite_all(lower_ld) || !finite_all(upper_l) || !finite_all(inverse_diagonal)) return false;
    if (!f64) return true;
    return finite_all(transpose_weights_f64) && finite_all(normal_bands_f64)
        && finite_all(ldlt_bands_f64) && finite_all(inverse_diagonal_f64);
}

std::size_t AxisPlan::storage_bytes() const noexcept {
    return sizeof(*this) + transpose_offsets.capacity() * 4 + transpose_indices.capacity() * 4
        + transpose_weights.capacity() * 4 + lower_ld.capacity() * 4 + upper_l.capacity() * 4
        + inverse_diagonal.capacity() * 4 + transpose_weights_f64.capacity() * 8
        + normal_bands_f64.capacity() * 8 + ldlt_bands_f64.capacity() * 8
        + inverse_diagonal_f64.capacity() * 8;
}

AxisPlan build_axis_plan(const AxisRequest &r, const CustomKernel &custom) {
    const trace::Range range(trace::Label::plan_build);
    const std::int32_t support = effective_support(r);
    if (r.kernel.kind == KernelKind::custom && !custom)
        throw std::invalid_argument("custom kernel callback is missing");
    const std::shared_ptr<const Geometry> geometry = geometry_cache().get(r, support);
    const SparseRows a = sample_operator(r, *geometry, custom);
    SparseRows t = transpose(a, r.source_size, r.destination_size);
    const std::int32_t n = r.destination_size;
    const std::int32_t w = std::min(2 * support - 1, n - 1);

    BandSystem g(n, w);
    g.form(t);
    BandSystem f = g;
    f.factor();
    const Conditioning cond = condition(g, f, n, w);
    const bool f64 = r.f64_mode == F64Mode::float64_only
        || (r.f64_mode == F64Mode::automatic && cond.rcond < kF64Threshold);

    AxisPlan p;
    p.source_size = r.source_size;
    p.destination_size = n;
    p.support = support;
    p.half_bandwidth = w;
    p.active_length = r.active_length;
    p.shift = r.shift;
    p.normal_rcond = cond.rcond;
    p.transpose_offsets = std::move(t.start);
    p.transpose_indices = std::move(t.column);
    p.transpose_weights.resize(t.value.size());
    std::transform(t.value.begin(), t.value.end(), p.transpose_weights.begin(),
                   [](double v) { return static_cast<float>(v); });
    const auto un = static_cast<std::size_t>(n);
    if (f64) {
        p.normal_inf_norm = cond.one_norm;
        p.transpose_weights_f64 = std::move(t.value);
        p.normal_bands_f64 = g.data();
        p.ldlt_bands_f64 = f.data();
        p.inverse_diagonal_f64.resize(un);
        for (std::int32_t i = 0; i < n; ++i)
            p.inverse_diagonal_f64[static_cast<std::size_t>(i)] = 1.0 / (f.at(0, i) + kEps);
    }
    p.lower_ld.assign(static_cast<std::size_t>(w) * un, 0.0F);
    p.upper_l.assign(static_cast<std::size_t>(w) * un, 0.0F);
    p.inverse_diagonal.resize(un);
    for (std::int32_t i = 0; i < n; ++i) {
        const double diag = f.at(0, i);
        p.inverse_diagonal[static_cast<std::size_t>(i)] = static_cast<float>(1.0 / (diag + kEps));
        for (std::int32_t d = 1, reach = std::min(w, n - i - 1); d <= reach; ++d) {
            const double l = f.at(d, i);
            p.upper_l[static_cast<std::size_t>(d - 1) * un + static_cast<std::size_t>(i)] = static_cast<float>(l);
            p.lower_ld[static_cast<std::size_t>(d - 1) * un + static_cast<std::size_t>(i + d)] =
                static_cast<float>(l * diag);
        }
    }
    if (!p.valid()) throw std::runtime_error("failed to build inverse axis plan");
    p.serial = next_plan_serial();
    return p;
}

std::uint64_t next_plan_serial() noexcept {
    static std::atomic<std::uint64_t> next_serial{1};
    return next_serial.fetch_add(1, std::memory_order_relaxed);
}

std::shared_ptr<const AxisPlan> import_axis_plan(AxisPlan plan) {
    // A plan made elsewhere (the reference planner, a caller's own arrays):
    // held to the same validity rules as this planner's output, then given
    // a serial of its own so device images can key on it.
    plan.serial = 0;
    if (!plan.valid()) throw std::invalid_argument("dsmvc: imported axis plan is invalid");
    plan.serial = next_plan_serial();
    return std::make_shared<const AxisPlan>(std::move(plan));
}

// ---------------------------------------------------------------------------
// Plan cache: single-flight LRU keyed by the exact bits of the request
// (the reference keys its plan LRU the same way, axis_plan.cpp:37-51).
// ---------------------------------------------------------------------------
namespace {

struct CacheKey {
    std::uint64_t word[10];
    bool operator==(const CacheKey &o) const { return std::memcmp(word, o.word, sizeof word) == 0; }
};

struct CacheKeyHash {
    std::size_t operator()(const CacheKey &k) const noexcept {
        std::uint64_t h = 0x9E3779B97F4A7C15ULL;
        for (const auto w : k.word) h = (h ^ w) * 0xBF58476D1CE4E5B9ULL + (h >> 31);
        return static_cast<std::size_t>(h);
    }
};

CacheKey key_of(const AxisRequest &r) {
    const bool bicubic = r.kernel.kind == KernelKind::bicubic;
    const bool lanczos = r.kernel.kind == KernelKind::lanczos;
    CacheKey k{};
    k.word[0] = static_cast<std::uint32_t>(r.source_size);
    k.word[1] = static_cast<std::uint32_t>(r.destination_size);
    k.word[2] = std::bit_cast<std::uint64_t>(r.active_length);
    k.word[3] = std::bit_cast<std::uint64_t>(r.shift);
    k.word[4] = static_cast<std::uint64_t>(r.kernel.kind) | (static_cast<std::uint64_t>(r.border) << 8)
        | (static_cast<std::uint64_t>(r.f64_mode) << 16);
    k.word[5] = lanczos ? static_cast<std::uint32_t>(r.kernel.taps) : 0U;
    k.word[6] = bicubic ? std::bit_cast<std::uint64_t>(r.kernel.b) : 0U;
    k.word[7] = bicubic ? std::bit_cast<std::uint64_t>(r.kernel.c) : 0U;
    k.word[8] = std::bit_cast<std::uint64_t>(r.kernel.blur);
    return k;
}

class PlanCache {
public:
    std::shared_ptr<const AxisPlan> get(const AxisRequest &r) {
        const CacheKey key = key_of(r);
        std::promise<std::shared_ptr<const AxisPlan>> promise;
        std::shared_future<std::shared_ptr<const AxisPlan>> pending;
        bool builder = false;
        {
            std::lock_guard lock(mutex_);
            auto it = map_.find(key);
            if (it != map_.end()) {
                ++hits_;
                lru_.splice(lru_.begin(), lru_, it->second.where);
                pending = it->second.value;
            } else {
                builder = true;
                pending = promise.get_future().share();
                lru_.push_front(key);
                map_.emplace(key, Entry{pending, lru_.begin(), 0});
            }
        }
        if (!builder) return pending.get();
        try {
            auto plan = std::make_shared<const AxisPlan>(build_axis_plan(r));
            {
                std::lock_guard lock(mutex_);
                ++builds_;
                auto it = map_.find(key);
                if (it != map_.end()) {
                    it->second.bytes = plan->storage_bytes();
                    bytes_ += it->second.bytes;
                }
                evict();
            }
            promise.set_value(plan);
            return plan;
        } catch (...) {
            {
                std::lock_guard lock(mutex_);
                auto it = map_.find(key);
                if (it != map_.end()) {
                    lru_.erase(it->second.where);
                    map_.erase(it);
                }
            }
            promise.set_exception(std::current_exception());
            throw;
        }
    }

    // A built entry (bytes set) counts as a hit, as in get(); an absent or
    // still-building one returns nullptr without registering anything.
    std::shared_ptr<const AxisPlan> find(const AxisRequest &r) {
        const CacheKey key = key_of(r);
        std::shared_future<std::shared_ptr<const AxisPlan>> ready;
        {
            std::lock_guard lock(mutex_);
            auto it = map_.find(key);
            if (it == map_.end() || it->second.bytes == 0) return nullptr;
            ++hits_;
            lru_.splice(lru_.begin(), lru_, it->second.where);
            ready = it->second.value;
        }
        return ready.get();
    }

    PlannerCacheStats stats() {
        std::lock_guard lock(mutex_);
        PlannerCacheStats s;
        s.plan_hits = hits_;
        s.plan_builds = builds_;
        s.plan_entries = map_.size();
        s.plan_resident_bytes = bytes_;
        return s;
    }

    void clear() {
        std::lock_guard lock(mutex_);
        map_.clear();
        lru_.clear();
        bytes_ = 0;
        hits_ = builds_ = 0;
    }

private:
    struct Entry {
        std::shared_future<std::shared_ptr<const AxisPlan>> value;
        std::list<CacheKey>::iterator where;
        std::size_t bytes;
    };

    void evict() {
        while ((map_.size() > kEntries || bytes_ > kBytes) && !lru_.empty()) {
            const CacheKey victim = lru_.back();
            auto it = map_.find(victim);
            if (it == map_.end() || it->second.bytes == 0) break;  // still building
            bytes_ -= it->second.bytes;
            lru_.pop_back();
            map_.erase(it);
        }
    }

    static constexpr std::size_t kEntries = 4096;
    static constexpr std::size_t kBytes = 512U << 20;
    std::mutex mutex_;
    std::unordered_map<CacheKey, Entry, CacheKeyHash> map_;
    std::list<CacheKey> lru_;
    std::size_t bytes_ = 0;
    std::uint64_t hits_ = 0;
    std::uint64_t builds_ = 0;
};

PlanCache &plan_cache() {
    static PlanCache cache;
    return cache;
}

} // namespace

std::shared_ptr<const AxisPlan> get_or_build_axis_plan(const AxisRequest &r,
                                                       const CustomKernel &custom) {
    // Custom kernels cannot be keyed (the callback has no identity the
    // cache could compare), so their plans are built on every request, as
    // the reference does (axis_plan.cpp:1367-1372).
    if (r.kernel.kind == KernelKind::custom)
        return std::make_shared<const AxisPlan>(build_axis_plan(r, custom));
    return plan_cache().get(r);
}

std::shared_ptr<const AxisPlan> find_axis_plan(const AxisRequest &r) { return plan_cache().find(r); }

PlannerCacheStats planner_cache_stats() {
    PlannerCacheStats s = plan_cache().stats();
    geometry_cache().stats(s);
    return s;
}

void clear_planner_caches() {
    plan_cache().clear();
    geometry_cache().clear();
}

} // namespace dsmvc
