// Device images of axis plans and chunk partitions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "device_types.hpp"
#include "engine.hpp"

namespace dsmvc::hip {

// Host-side packing of one AxisPlan into the layouts the kernels read:
// CSR arrays as-is, coefficient rows (see device_types.hpp), and the FP64
// route's rows.  Pure host code, unit-testable without a device.
struct PackedPlan {
    std::int32_t n = 0, m = 0, w = 0, nnz = 0;
    std::int32_t coef_stride = 16;
    bool zform = false;
    std::vector<float> coef;         // [n * coef_stride]
    std::vector<double> coef64;      // [n * coef_stride]
    std::vector<double> wt64;        // [nnz]
};
PackedPlan pack_plan(const AxisPlan &plan);

// A chunk partition of a plan for the chunked kernels, built on the host in
// FP64 from the F32 coefficients (transfer matrices stored F32).
struct ChunkPlan {
    std::int32_t chunks = 0;
    std::int32_t length = 0;
    bool ring_feasible = false;      // rows_chunked can stage every CSR row
    std::vector<float> tf, tb;       // [chunks][W][W]
    std::vector<std::int32_t> win;   // [chunks][2]
};
// Returns false when the plan cannot be chunked with `chunks` chunks.
bool build_chunk_plan(const AxisPlan &plan, int chunks, ChunkPlan &out);
// The chunk count the executor uses for an axis of length n, or 0 for the
// sequential route.
int choose_chunks(const AxisPlan &plan, bool rows);

// One device allocation holding a plan's image (and its chunk partitions).
class DevicePlan {
public:
    DevicePlan(const AxisPlan &plan, int device);
    ~DevicePlan();
    DevicePlan(const DevicePlan &) = delete;
    DevicePlan &operator=(const DevicePlan &) = delete;

    [[nodiscard]] const DevAxis &axis() const { return axis_; }
    [[nodiscard]] const AxisPlan &host() const { return *plan_; }
    // Chunk partition on the device; nullptr when the plan is not chunkable
    // with that count.  Built lazily, cached.
    const DevChunks *chunks(int count, bool rows);
    [[nodiscard]] std::size_t device_bytes() const { return bytes_; }

private:
    struct ChunkImage {
        DevChunks dev{};
        void *memory = nullptr;
        bool ok = false;
        bool ring_feasible = false;
    };
    std::shared_ptr<const AxisPlan> keep_;
    const AxisPlan *plan_;
    int device_;
    void *memory_ = nullptr;
    std::size_t bytes_ = 0;
    DevAxis axis_{};
    std::mutex mutex_;
    std::map<int, ChunkImage> chunk_images_;
};

} // namespace dsmvc::hip
