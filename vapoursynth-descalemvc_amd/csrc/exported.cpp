// GPU frames exported by another API: the executor's *_exported entries
// (reference include/dsmvc/engine.hpp:306-326, 414-442 and
// src/cuda/cuda_executor.cpp:1180-1300, 3640-3900 for its CUDA executor).
// A frame producer (VapourSynth's GPU core) exports each allocation as a
// POSIX file descriptor; HIP imports it as external memory
// (hipImportExternalMemory + hipExternalMemoryGetMappedBuffer) once per
// memory_id, and every plane of it is addressed by offset.  The producer's
// timeline semaphore, when imported, is waited for on the device.
#include "exported.hpp"

#include <unistd.h>

#include <functional>
#include <iterator>
#include <stdexcept>
#include <string>
#include <utility>

#include "trace.hpp"

namespace dsmvc {
namespace {

// VkExternalMemoryHandleTypeFlagBits / VkExternalSemaphoreHandleTypeFlagBits
// OPAQUE_FD: the handle is a POSIX file descriptor.
constexpr int kOpaqueFd = 0x1;

void close_handle(std::intptr_t handle, int type) noexcept {
    if (type == kOpaqueFd && handle >= 0) (void)::close(static_cast<int>(handle));
}

void check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
    }
}

// Restores the calling thread's current device.
struct DeviceGuard {
    int previous = -1;
    explicit DeviceGuard(int device) {
        if (hipGetDevice(&previous) != hipSuccess) previous = -1;
        check(hipSetDevice(device), "hipSetDevice failed");
    }
    ~DeviceGuard() {
        if (previous >= 0) (void)hipSetDevice(previous);
    }
};

std::size_t sample_bytes(SampleKind kind) {
    return kind == SampleKind::float32 ? 4 : (kind == SampleKind::u16 ? 2 : 1);
}

struct Resolved {
    std::shared_ptr<ImportedAllocation> allocation;
    unsigned char *pointer = nullptr;
};

// Imports (or finds) the plane's allocation after checking that `rows` rows
// of `row_bytes` at the plane's pitch lie inside the plane and the
// allocation.  The handle is consumed on every path.
Resolved resolve(HipExecutor::Imports &imports, const ExportedPlane &plane, std::size_t row_bytes,
                 std::size_t rows, std::size_t element) {
    auto reject = [&](const char *message) {
        close_handle(plane.handle, plane.handle_type);
        throw std::invalid_argument(message);
    };
    if (plane.pitch_bytes <= 0 || plane.pitch_bytes % static_cast<std::ptrdiff_t>(element) != 0
        || row_bytes == 0 || rows == 0 || static_cast<std::size_t>(plane.pitch_bytes) < row_bytes)
        reject("dsmvc: invalid exported GPU plane geometry");
    const std::size_t span = (rows - 1) * static_cast<std::size_t>(plane.pitch_bytes) + row_bytes;
    if (plane.offset > plane.memory_size || plane.plane_bytes > plane.memory_size - plane.offset
        || span > plane.plane_bytes)
        reject("dsmvc: exported GPU plane lies outside its allocation");
    Resolved r;
    r.allocation = imports.acquire(plane);
    r.pointer = r.allocation->base() + plane.offset;
    return r;
}

// Both planes of a call; the output's handle is closed when the input fails.
std::pair<Resolved, Resolved> resolve_pair(HipExecutor::Imports &imports, const ExportedPlane &in,
                                           std::size_t in_row_bytes, std::size_t in_rows,
                                           const ExportedPlane &out, std::size_t out_row_bytes,
                                           std::size_t out_rows, std::size_t element) {
    Resolved source;
    try {
        source = resolve(imports, in, in_row_bytes, in_rows, element);
    } catch (...) {
        close_handle(out.handle, out.handle_type);
        throw;
    }
    return {std::move(source), resolve(imports, out, out_row_bytes, out_rows, element)};
}

struct Release {
    std::function<void()> run;
    ~Release() {
        if (run) run();
    }
};

} // namespace

void close_exported_plane(const ExportedPlane &plane) noexcept {
    close_handle(plane.handle, plane.handle_type);
}

ImportedAllocation::ImportedAllocation(int device, const ExportedPlane &plane) : device_(device) {
    if (plane.handle_type != kOpaqueFd || plane.handle < 0) {
        close_handle(plane.handle, plane.handle_type);
        throw std::runtime_error("dsmvc: the producer exports GPU memory in a form HIP cannot import");
    }
    DeviceGuard guard(device_);
    hipExternalMemoryHandleDesc desc{};
    desc.type = hipExternalMemoryHandleTypeOpaqueFd;
    desc.handle.fd = static_cast<int>(plane.handle);
    desc.size = plane.memory_size;
    hipError_t status = hipImportExternalMemory(&memory_, &desc);
    if (status != hipSuccess) {
        // A dedicated allocation must be imported as one; the handle is only
        // consumed by a successful import, so it can be tried again.
        (void)hipGetLastError();
        desc.flags = hipExternalMemoryDedicated;
        status = hipImportExternalMemory(&memory_, &desc);
    }
    if (status != hipSuccess) {
        (void)hipGetLastError();
        memory_ = nullptr;
        close_handle(plane.handle, plane.handle_type);
        throw std::runtime_error(std::string("dsmvc: hipImportExternalMemory failed: ")
                                 + hipGetErrorString(status));
    }
    hipExternalMemoryBufferDesc buffer{};
    buffer.size = plane.memory_size;
    void *mapped = nullptr;
    status = hipExternalMemoryGetMappedBuffer(&mapped, memory_, &buffer);
    if (status != hipSuccess) {
        (void)hipGetLastError();
        (void)hipDestroyExternalMemory(memory_);
        memory_ = nullptr;
        throw std::runtime_error(std::string("dsmvc: hipExternalMemoryGetMappedBuffer failed: ")
                                 + hipGetErrorString(status));
    }
    base_ = static_cast<unsigned char *>(mapped);
}

ImportedAllocation::~ImportedAllocation() {
    int previous = -1;
    if (hipGetDevice(&previous) != hipSuccess) previous = -1;
    (void)hipSetDevice(device_);
    if (base_) (void)hipFree(base_);
    if (memory_) (void)hipDestroyExternalMemory(memory_);
    (void)hipGetLastError();
    if (previous >= 0) (void)hipSetDevice(previous);
}

HipExecutor::Imports::~Imports() {
    entries_.clear();
    int previous = -1;
    if (hipGetDevice(&previous) != hipSuccess) previous = -1;
    (void)hipSetDevice(device_);
    for (auto &entry : semaphores_)
        if (entry.second) (void)hipDestroyExternalSemaphore(entry.second);
    if (previous >= 0) (void)hipSetDevice(previous);
}

std::shared_ptr<ImportedAllocation> HipExecutor::Imports::acquire(const ExportedPlane &plane) {
    {
        std::lock_guard lock(mutex_);
        if (auto it = entries_.find(plane.memory_id); it != entries_.end()) {
            lru_.splice(lru_.end(), lru_, it->second.lru);
            ++stats_.hits;
            close_handle(plane.handle, plane.handle_type);
            return it->second.allocation;
        }
    }
    // Imported outside the lock; the constructor consumes or closes the handle.
    auto allocation = std::make_shared<ImportedAllocation>(device_, plane);
    std::lock_guard lock(mutex_);
    if (auto it = entries_.find(plane.memory_id); it != entries_.end())
        return it->second.allocation;   // another thread was first; ours is freed
    ++stats_.imports;
    lru_.push_back(plane.memory_id);
    entries_.emplace(plane.memory_id, Entry{allocation, std::prev(lru_.end())});
    // Bounded: imports of allocations the producer has freed would keep that
    // memory alive.  Entries a call still holds are skipped.
    for (auto k = lru_.begin(); entries_.size() > kEntries && k != lru_.end();) {
        auto it = entries_.find(*k);
        if (it->second.allocation.use_count() > 1) {
            ++k;
            continue;
        }
        entries_.erase(it);
        k = lru_.erase(k);
        ++stats_.evictions;
    }
    return allocation;
}

bool HipExecutor::Imports::import_semaphore(std::uint64_t semaphore, int handle_type,
                                            std::intptr_t handle) {
    hipExternalSemaphore_t imported = nullptr;
    bool consumed = false;
    if (handle_type == kOpaqueFd && handle >= 0 && semaphore != 0) {
        try {
            DeviceGuard guard(device_);
            hipExternalSemaphoreHandleDesc desc{};
            desc.type = hipExternalSemaphoreHandleTypeTimelineSemaphoreFd;
            desc.handle.fd = static_cast<int>(handle);
            if (hipImportExternalSemaphore(&imported, &desc) == hipSuccess) {
                consumed = true;
            } else {
                (void)hipGetLastError();
                imported = nullptr;
            }
        } catch (const std::exception &) {
            imported = nullptr;
        }
    }
    if (!consumed) close_handle(handle, handle_type);
    std::lock_guard lock(mutex_);
    const auto [entry, inserted] = semaphores_.emplace(semaphore, imported);
    if (!inserted && imported) {
        // Another thread imported it first.
        int previous = -1;
        if (hipGetDevice(&previous) != hipSuccess) previous = -1;
        (void)hipSetDevice(device_);
        (void)hipDestroyExternalSemaphore(imported);
        if (previous >= 0) (void)hipSetDevice(previous);
    }
    return entry->second != nullptr;
}

hipExternalSemaphore_t HipExecutor::Imports::semaphore(std::uint64_t semaphore) {
    std::lock_guard lock(mutex_);
    auto it = semaphores_.find(semaphore);
    return it == semaphores_.end() ? nullptr : it->second;
}

HipExecutor::ImportStats HipExecutor::Imports::stats() {
    std::lock_guard lock(mutex_);
    ImportStats s = stats_;
    s.entries = entries_.size();
    return s;
}

HipExecutor::Imports &HipExecutor::imports() {
    std::lock_guard lock(imports_mutex_);
    if (!imports_) imports_ = std::make_unique<Imports>(device_);
    return *imports_;
}

bool HipExecutor::import_exported_semaphore(std::uint64_t semaphore, int handle_type,
                                            std::intptr_t handle) {
    return imports().import_semaphore(semaphore, handle_type, handle);
}

bool HipExecutor::has_exported_semaphore(std::uint64_t semaphore) {
    return imports().semaphore(semaphore) != nullptr;
}

HipExecutor::ImportStats HipExecutor::import_stats() { return imports().stats(); }

namespace {
// Queues the wait for an input frame's producer pair ahead of the stream's
// reads when its semaphore was imported; otherwise the caller has waited.
void wait_ready(HipExecutor::Imports &imports, const ExportedPlane &input, hipStream_t stream) {
    if (!input.ready_semaphore) return;
    hipExternalSemaphore_t semaphore = imports.semaphore(input.ready_semaphore);
    if (!semaphore) return;
    hipExternalSemaphoreWaitParams params{};
    params.params.fence.value = input.ready_value;
    check(hipWaitExternalSemaphoresAsync(&semaphore, &params, 1, stream),
          "dsmvc: hipWaitExternalSemaphoresAsync failed");
}
} // namespace

void HipExecutor::inverse_2d_exported(const AxisPlan &h, const AxisPlan &v, SampleKind kind,
                                      const ExportedPlane &input, const ExportedPlane &output,
                                      const IntegerConversion *conversion) {
    const trace::Range range(trace::Label::execute_exported);
    const bool integer = kind != SampleKind::float32;
    if (!h.valid() || !v.valid() || (integer && !conversion)) {
        close_exported_plane(input);
        close_exported_plane(output);
        throw std::invalid_argument(integer ? "dsmvc: invalid integer 2D executor arguments"
                                            : "dsmvc: invalid 2D executor arguments");
    }
    const std::size_t e = sample_bytes(kind);
    const auto pair = resolve_pair(imports(), input, static_cast<std::size_t>(h.source_size) * e,
                                   static_cast<std::size_t>(v.source_size), output,
                                   static_cast<std::size_t>(h.destination_size) * e,
                                   static_cast<std::size_t>(v.destination_size), e);
    Slot *slot = acquire_slot();
    Release release{[this, slot] { release_slot(slot); }};
    {
        DeviceGuard guard(device_);
        wait_ready(imports(), input, slot->stream);
    }
    const PlaneBatch batch{pair.first.pointer, input.pitch_bytes / static_cast<std::ptrdiff_t>(e), 0,
                           pair.second.pointer, output.pitch_bytes / static_cast<std::ptrdiff_t>(e),
                           0, 1};
    inverse_2d_batch(h, v, kind, batch, conversion, slot->stream, false);
}

void HipExecutor::inverse_rows_exported(const AxisPlan &plan, const ExportedPlane &input,
                                        const ExportedPlane &output, std::int32_t row_count) {
    const trace::Range range(trace::Label::execute_exported);
    if (!plan.valid() || row_count <= 0) {
        close_exported_plane(input);
        close_exported_plane(output);
        if (row_count == 0 && plan.valid()) return;
        throw std::invalid_argument("dsmvc: invalid row executor arguments");
    }
    const auto rows = static_cast<std::size_t>(row_count);
    const auto pair = resolve_pair(imports(), input, static_cast<std::size_t>(plan.source_size) * 4,
                                   rows, output, static_cast<std::size_t>(plan.destination_size) * 4,
                                   rows, 4);
    Slot *slot = acquire_slot();
    Release release{[this, slot] { release_slot(slot); }};
    {
        DeviceGuard guard(device_);
        wait_ready(imports(), input, slot->stream);
    }
    inverse_rows(plan, reinterpret_cast<const float *>(pair.first.pointer), input.pitch_bytes / 4,
                 reinterpret_cast<float *>(pair.second.pointer), output.pitch_bytes / 4, row_count,
                 slot->stream, false);
}

void HipExecutor::inverse_columns_exported(const AxisPlan &plan, const ExportedPlane &input,
                                           const ExportedPlane &output, std::int32_t column_count) {
    const trace::Range range(trace::Label::execute_exported);
    if (!plan.valid() || column_count <= 0) {
        close_exported_plane(input);
        close_exported_plane(output);
        if (column_count == 0 && plan.valid()) return;
        throw std::invalid_argument("dsmvc: invalid column executor arguments");
    }
    const auto columns = static_cast<std::size_t>(column_count) * 4;
    const auto pair = resolve_pair(imports(), input, columns,
                                   static_cast<std::size_t>(plan.source_size), output, columns,
                                   static_cast<std::size_t>(plan.destination_size), 4);
    Slot *slot = acquire_slot();
    Release release{[this, slot] { release_slot(slot); }};
    {
        DeviceGuard guard(device_);
        wait_ready(imports(), input, slot->stream);
    }
    inverse_columns(plan, reinterpret_cast<const float *>(pair.first.pointer), input.pitch_bytes / 4,
                    reinterpret_cast<float *>(pair.second.pointer), output.pitch_bytes / 4,
                    column_count, slot->stream, false);
}

} // namespace dsmvc
