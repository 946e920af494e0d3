// Test infrastructure (not part of the library): device memory exported as a
// POSIX file descriptor, standing in for a frame that VapourSynth's GPU core
// exports to a foreign API.  Allocated with the HIP virtual memory API
// (hipMemCreate with a POSIX-descriptor handle type), mapped into this
// process for filling and reading, and exported once per call of
// xp_export, as the producer exports every frame anew.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

extern "C" {

// Allocates at least `bytes` on `device`; *size receives the rounded size.
int xp_create(int device, std::size_t bytes, std::uint64_t *handle, void **ptr, std::size_t *size) {
    hipMemAllocationProp prop{};
    prop.type = hipMemAllocationTypePinned;
    prop.requestedHandleType = hipMemHandleTypePosixFileDescriptor;
    prop.location.type = hipMemLocationTypeDevice;
    prop.location.id = device;
    std::size_t granularity = 0;
    hipError_t e = hipMemGetAllocationGranularity(&granularity, &prop, hipMemAllocationGranularityMinimum);
    if (e != hipSuccess) return static_cast<int>(e);
    const std::size_t rounded = (bytes + granularity - 1) / granularity * granularity;
    hipMemGenericAllocationHandle_t h{};
    if ((e = hipMemCreate(&h, rounded, &prop, 0)) != hipSuccess) return static_cast<int>(e);
    void *va = nullptr;
    if ((e = hipMemAddressReserve(&va, rounded, granularity, nullptr, 0)) != hipSuccess) {
        (void)hipMemRelease(h);
        return static_cast<int>(e);
    }
    if ((e = hipMemMap(va, rounded, 0, h, 0)) != hipSuccess) {
        (void)hipMemAddressFree(va, rounded);
        (void)hipMemRelease(h);
        return static_cast<int>(e);
    }
    hipMemAccessDesc access{};
    access.location = prop.location;
    access.flags = hipMemAccessFlagsProtReadWrite;
    if ((e = hipMemSetAccess(va, rounded, &access, 1)) != hipSuccess) {
        (void)hipMemUnmap(va, rounded);
        (void)hipMemAddressFree(va, rounded);
        (void)hipMemRelease(h);
        return static_cast<int>(e);
    }
    *handle = reinterpret_cast<std::uint64_t>(h);
    *ptr = va;
    *size = rounded;
    return 0;
}

// A new file descriptor for the allocation.
int xp_export(std::uint64_t handle, int *fd) {
    return static_cast<int>(hipMemExportToShareableHandle(
        fd, reinterpret_cast<hipMemGenericAllocationHandle_t>(handle),
        hipMemHandleTypePosixFileDescriptor, 0));
}

int xp_copy_in(void *dst, const void *host, std::size_t bytes) {
    return static_cast<int>(hipMemcpy(dst, host, bytes, hipMemcpyHostToDevice));
}

int xp_copy_out(void *host, const void *src, std::size_t bytes) {
    return static_cast<int>(hipMemcpy(host, src, bytes, hipMemcpyDeviceToHost));
}

int xp_destroy(std::uint64_t handle, void *ptr, std::size_t size) {
    (void)hipDeviceSynchronize();
    hipError_t e = hipMemUnmap(ptr, size);
    if (e == hipSuccess) e = hipMemAddressFree(ptr, size);
    if (e == hipSuccess) e = hipMemRelease(reinterpret_cast<hipMemGenericAllocationHandle_t>(handle));
    return static_cast<int>(e);
}

} // extern "C"
