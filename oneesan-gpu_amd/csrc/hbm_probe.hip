// hbm_probe.hip -- what plain streaming kernels reach on this GPU's HBM, the
// practical ceiling the cell kernels are compared with.
//
//   ofp_hbm_probe [GiB per buffer, default 64]
//   ofp_hbm_probe ws              working sets of 16 MiB .. 2 GiB, repeated passes
//
// Three grid-stride kernels over one buffer of B bytes (and a second for the
// copy): read-only (B bytes moved), copy (2B) and in-place read-modify-write
// (2B), with 8-byte (uint2) and 16-byte (uint4) accesses, at 2048, 8192 and
// 32768 workgroups of 256 threads; each timed with HIP events, best of two.
// `ws` times 20 back-to-back read and read-modify-write passes (16-byte
// accesses) over one working set, for sizes around the 256 MB Infinity Cache:
// what a kernel that keeps re-sweeping a cache-resident block would see.
// Prints one JSON object per measurement.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

template <typename T>
__global__ void read_kernel(const T* __restrict__ a, size_t n, unsigned* sink) {
    unsigned acc = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        acc ^= a[i].x;
    if (acc == 0x9e3779b9u) sink[0] = acc;  // keeps the loads; the buffer never holds this pattern
}

template <typename T>
__global__ void copy_kernel(const T* __restrict__ a, T* __restrict__ b, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

template <typename T>
__global__ void rmw_kernel(T* __restrict__ a, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        T v = a[i];
        v.x += 1u;
        a[i] = v;
    }
}

bool ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    fprintf(stderr, "ofp_hbm_probe: %s: %s\n", what, hipGetErrorString(e));
    return false;
}

template <typename T>
bool run(const char* kind, char* a, char* b, unsigned* sink, size_t bytes, int grid, hipEvent_t e0,
         hipEvent_t e1) {
    const size_t n = bytes / sizeof(T);
    float best = 1e30f;
    double moved = 0;
    for (int rep = 0; rep < 2; ++rep) {
        if (!ok(hipEventRecord(e0), "event")) return false;
        if (kind[0] == 'r' && kind[1] == 'e') {
            hipLaunchKernelGGL(read_kernel<T>, dim3(grid), dim3(256), 0, 0, (const T*)a, n, sink);
            moved = (double)bytes;
        } else if (kind[0] == 'c') {
            hipLaunchKernelGGL(copy_kernel<T>, dim3(grid), dim3(256), 0, 0, (const T*)a, (T*)b, n);
            moved = 2.0 * bytes;
        } else {
            hipLaunchKernelGGL(rmw_kernel<T>, dim3(grid), dim3(256), 0, 0, (T*)a, n);
            moved = 2.0 * bytes;
        }
        if (!ok(hipGetLastError(), "launch") || !ok(hipEventRecord(e1), "event") ||
            !ok(hipEventSynchronize(e1), "sync"))
            return false;
        float ms = 0;
        if (!ok(hipEventElapsedTime(&ms, e0, e1), "elapsed")) return false;
        if (ms < best) best = ms;
    }
    printf("{\"kernel\": \"%s\", \"access_bytes\": %zu, \"workgroups\": %d, \"bytes_moved\": %.0f, "
           "\"ms\": %.3f, \"GBps\": %.1f}\n",
           kind, sizeof(T), grid, moved, best, moved / best / 1e6);
    fflush(stdout);
    return true;
}

}  // namespace

// Repeated passes over working sets of growing size.
int working_sets() {
    const size_t mib[] = {16, 32, 64, 128, 192, 256, 384, 512, 1024, 2048};
    const int passes = 20;
    char* a = nullptr;
    unsigned* sink = nullptr;
    hipEvent_t e0, e1;
    const size_t maxb = (size_t)2048 << 20;
    if (!ok(hipMalloc(&a, maxb), "hipMalloc") || !ok(hipMalloc(&sink, 64), "hipMalloc") ||
        !ok(hipMemset(a, 1, maxb), "memset") || !ok(hipEventCreate(&e0), "event") ||
        !ok(hipEventCreate(&e1), "event") || !ok(hipDeviceSynchronize(), "sync"))
        return 2;
    for (size_t m : mib) {
        const size_t bytes = m << 20, n = bytes / sizeof(uint4);
        const size_t want = n / 256;
        const int grid = (int)(want < 32768 ? (want > 256 ? want : 256) : 32768);
        for (int kind = 0; kind < 2; ++kind) {
            float best = 1e30f;
            for (int rep = 0; rep < 2; ++rep) {
                // one untimed pass brings the set into the caches
                if (kind == 0) hipLaunchKernelGGL(read_kernel<uint4>, dim3(grid), dim3(256), 0, 0, (const uint4*)a, n, sink);
                else hipLaunchKernelGGL(rmw_kernel<uint4>, dim3(grid), dim3(256), 0, 0, (uint4*)a, n);
                if (!ok(hipEventRecord(e0), "event")) return 3;
                for (int i = 0; i < passes; ++i) {
                    if (kind == 0) hipLaunchKernelGGL(read_kernel<uint4>, dim3(grid), dim3(256), 0, 0, (const uint4*)a, n, sink);
                    else hipLaunchKernelGGL(rmw_kernel<uint4>, dim3(grid), dim3(256), 0, 0, (uint4*)a, n);
                }
                if (!ok(hipGetLastError(), "launch") || !ok(hipEventRecord(e1), "event") ||
                    !ok(hipEventSynchronize(e1), "sync"))
                    return 3;
                float ms = 0;
                if (!ok(hipEventElapsedTime(&ms, e0, e1), "elapsed")) return 3;
                if (ms < best) best = ms;
            }
            const double moved = (kind == 0 ? 1.0 : 2.0) * (double)bytes * passes;
            printf("{\"kernel\": \"%s\", \"working_set_mib\": %zu, \"passes\": %d, \"workgroups\": %d, "
                   "\"bytes_moved\": %.0f, \"ms\": %.3f, \"GBps\": %.1f}\n",
                   kind == 0 ? "read" : "rmw", m, passes, grid, moved, best, moved / best / 1e6);
            fflush(stdout);
        }
    }
    (void)hipFree(a);
    (void)hipFree(sink);
    return 0;
}

int main(int argc, char** argv) {
    if (argc > 1 && argv[1][0] == 'w' && argv[1][1] == 's') return working_sets();
    const double gib = argc > 1 ? atof(argv[1]) : 64.0;
    if (!(gib >= 1.0 && gib <= 120.0)) {
        fprintf(stderr, "ofp_hbm_probe: buffer size must be 1..120 GiB\n");
        return 1;
    }
    const size_t bytes = ((size_t)(gib * (double)(1ull << 30))) & ~(size_t)4095;
    char *a = nullptr, *b = nullptr;
    unsigned* sink = nullptr;
    hipEvent_t e0, e1;
    if (!ok(hipMalloc(&a, bytes), "hipMalloc") || !ok(hipMalloc(&b, bytes), "hipMalloc") ||
        !ok(hipMalloc(&sink, 64), "hipMalloc") || !ok(hipMemset(a, 1, bytes), "memset") ||
        !ok(hipMemset(b, 0, bytes), "memset") || !ok(hipEventCreate(&e0), "event") ||
        !ok(hipEventCreate(&e1), "event") || !ok(hipDeviceSynchronize(), "sync"))
        return 2;
    const int grids[3] = {2048, 8192, 32768};
    for (int g : grids) {
        if (!run<uint2>("read", a, b, sink, bytes, g, e0, e1) || !run<uint4>("read", a, b, sink, bytes, g, e0, e1) ||
            !run<uint2>("copy", a, b, sink, bytes, g, e0, e1) || !run<uint4>("copy", a, b, sink, bytes, g, e0, e1) ||
            !run<uint2>("rmw", a, b, sink, bytes, g, e0, e1) || !run<uint4>("rmw", a, b, sink, bytes, g, e0, e1))
            return 3;
    }
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(sink);
    return 0;
}
