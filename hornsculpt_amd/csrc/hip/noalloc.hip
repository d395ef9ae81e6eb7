// Dense complex LU solve that allocates no device memory.
//
// The factorization needs device memory beyond the caller's matrix and
// right-hand sides: the pivot rows (n int32) and the singular-pivot status
// word. Here they live in one file-scope device array (`scratch`) instead of
// hipMalloc'd buffers, so the entry point can serve callers that may not
// allocate (code under stream capture, tight serving loops, harnesses that
// forbid hipMalloc). Every GPU the library runs on holds its own copy of the
// array until the process ends; its device address is resolved once per
// device with hipGetSymbolAddress and only that address is handed to kernels
// and copies. Calls that overlap on one GPU share the array and race.
#include "../solver.hpp"
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <array>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>

namespace {
// Largest n the scratch array holds pivots for.
constexpr std::uint32_t kNoAllocMaxOrder = 65536;
// Layout: status word at offset 0 (256-byte slot), pivots from offset 256.
constexpr std::size_t kStatusBytes = 256;
constexpr std::size_t kScratchBytes = kStatusBytes + sizeof(std::int32_t) * kNoAllocMaxOrder;
} // namespace

// 256 KiB + 256 B per GPU, held until the process ends.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace hsbem::dev {

namespace {

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
std::array<unsigned char *, kMaxDevices> scratch_address{};

// Device address of this device's copy of `scratch`; the device must be current.
unsigned char *device_scratch(int device) {
    if (device < 0 || device >= kMaxDevices)
        throw std::invalid_argument("device ordinal out of range for the no-allocation scratch");
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_address[device]) {
        void *ptr = nullptr;
        hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "scratch symbol address");
        if (!ptr)
            throw std::runtime_error("scratch symbol has no device address");
        scratch_address[device] = static_cast<unsigned char *>(ptr);
    }
    return scratch_address[device];
}

// Clears the status word with a kernel rather than a memset node, so a
// captured stream records no memset.
__global__ void clear_status(int *status) { *status = 0; }

} // namespace

std::uint32_t noalloc_lu_max_order() { return kNoAllocMaxOrder; }

void dense_lu_solve_noalloc(int device, hipStream_t stream, std::uint32_t n, double *A, std::uint32_t nrhs,
                            double *B) {
    if (!A || !B || !n || !nrhs)
        throw std::invalid_argument("invalid LU arguments");
    if (n > kNoAllocMaxOrder)
        throw std::invalid_argument("LU order " + std::to_string(n) + " exceeds the no-allocation limit of " +
                                    std::to_string(kNoAllocMaxOrder));
    hip_check(hipSetDevice(device), "set device");
    unsigned char *base = device_scratch(device);
    int *status = reinterpret_cast<int *>(base);
    auto *pivots = reinterpret_cast<std::int32_t *>(base + kStatusBytes);
    hipLaunchKernelGGL(clear_status, dim3(1), dim3(1), 0, stream, status);
    hip_check(hipGetLastError(), "status clear");
    lu_factor(A, n, pivots, status, nullptr, stream);
    int code = 0;
    hip_check(hipMemcpyAsync(&code, status, sizeof(int), hipMemcpyDeviceToHost, stream), "status download");
    hip_check(hipStreamSynchronize(stream), "LU sync");
    if (code)
        throw std::runtime_error("LU factorization failed at pivot " + std::to_string(code));
    lu_solve(A, n, pivots, B, nrhs, stream);
    hip_check(hipStreamSynchronize(stream), "LU solve sync");
}

} // namespace hsbem::dev
