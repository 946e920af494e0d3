// Imports of GPU frame memory exported by another API, for the executor's
// *_exported entries (exported.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <list>
#include <memory>
#include <mutex>
#include <unordered_map>

#include "engine.hpp"
#include "executor.hpp"

namespace dsmvc {

// One exported allocation imported into HIP and mapped: the planes of its
// frames are addressed by offset.  Freed when the cache has dropped it and
// the last call using it has finished.
class ImportedAllocation {
public:
    ImportedAllocation(int device, const ExportedPlane &plane);
    ~ImportedAllocation();
    ImportedAllocation(const ImportedAllocation &) = delete;
    ImportedAllocation &operator=(const ImportedAllocation &) = delete;
    [[nodiscard]] unsigned char *base() const noexcept { return base_; }

private:
    int device_;
    hipExternalMemory_t memory_ = nullptr;
    unsigned char *base_ = nullptr;
};

class HipExecutor::Imports {
public:
    explicit Imports(int device) : device_(device) {}
    ~Imports();
    Imports(const Imports &) = delete;
    Imports &operator=(const Imports &) = delete;

    // The allocation of `plane`, imported on first use (keyed by memory_id,
    // which the producer never reuses).  Every export is a new handle: a hit
    // closes it.
    std::shared_ptr<ImportedAllocation> acquire(const ExportedPlane &plane);
    bool import_semaphore(std::uint64_t semaphore, int handle_type, std::intptr_t handle);
    // The imported semaphore for `semaphore`, or nullptr.
    hipExternalSemaphore_t semaphore(std::uint64_t semaphore);
    HipExecutor::ImportStats stats();

private:
    static constexpr std::size_t kEntries = 64;
    int device_;
    std::mutex mutex_;
    struct Entry {
        std::shared_ptr<ImportedAllocation> allocation;
        std::list<std::uint64_t>::iterator lru;
    };
    std::unordered_map<std::uint64_t, Entry> entries_;
    std::list<std::uint64_t> lru_;   // least recent first
    std::unordered_map<std::uint64_t, hipExternalSemaphore_t> semaphores_;
    HipExecutor::ImportStats stats_{};
};

} // namespace dsmvc
