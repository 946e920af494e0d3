// roctx ranges over the executor's host-side stages, for rocprofv3
// --marker-trace (the reference marks the same stages of its CUDA executor
// with an NVTX domain, src/cuda/nvtx.hpp:11-45, used in
// src/cuda/cuda_executor.cpp:2552-2973).  Every range name starts with
// "dsmvc.hip: " so a trace can be filtered to this library.  A range is a
// push on construction and a pop on destruction, on the calling thread;
// without a profiler attached roctx calls return at once.
#pragma once

#include <rocprofiler-sdk-roctx/roctx.h>

#include <array>
#include <cstdint>

namespace dsmvc::trace {

enum class Label : std::uint8_t {
    execute_2d,
    execute_planes,
    execute_rows,
    execute_columns,
    execute_host_frame,
    execute_exported,
    plan_build,
    geometry_build,
    plan_cache_lookup,
    plan_upload,
    plan_evict,
    window_build,
    slot_wait,
    buffer_reserve,
    table_upload,
    horizontal_stage,
    vertical_stage,
    fused_stage,
    operator_build,
    stream_synchronize,
    host_pack,
    source_upload,
    destination_download,
    host_unpack,
    batch_setup,
    batch_run,
    server_submit,
    server_round,
    count,
};

inline constexpr std::array<const char *, static_cast<std::size_t>(Label::count)> kNames{
    "dsmvc.hip: execute 2d",
    "dsmvc.hip: execute planes",
    "dsmvc.hip: execute rows",
    "dsmvc.hip: execute columns",
    "dsmvc.hip: execute host frame",
    "dsmvc.hip: execute exported frame",
    "dsmvc.hip: plan build",
    "dsmvc.hip: geometry build",
    "dsmvc.hip: plan cache lookup",
    "dsmvc.hip: plan upload",
    "dsmvc.hip: plan evict",
    "dsmvc.hip: window build",
    "dsmvc.hip: slot wait",
    "dsmvc.hip: buffer reserve",
    "dsmvc.hip: table upload",
    "dsmvc.hip: horizontal stage",
    "dsmvc.hip: vertical stage",
    "dsmvc.hip: fused 2d stage",
    "dsmvc.hip: operator build",
    "dsmvc.hip: stream synchronize",
    "dsmvc.hip: host pack",
    "dsmvc.hip: source upload",
    "dsmvc.hip: destination download",
    "dsmvc.hip: host unpack",
    "dsmvc.hip: batch setup",
    "dsmvc.hip: batch run",
    "dsmvc.hip: server submit",
    "dsmvc.hip: server round",
};

class Range {
public:
    explicit Range(Label label) noexcept { (void)roctxRangePushA(kNames[static_cast<std::size_t>(label)]); }
    ~Range() { (void)roctxRangePop(); }
    Range(const Range &) = delete;
    Range &operator=(const Range &) = delete;
};

// A point event: a plan cache hit or miss.
inline void mark(const char *what) noexcept { roctxMarkA(what); }

} // namespace dsmvc::trace
