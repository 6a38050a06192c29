// Atomic throughput of the MI355X for the operations the raster path relies
// on, measured with HIP events (the roofline's atomic half):
//
//   lds_min_u64     64-bit LDS atomicMin without return (ds_min_u64), every
//                   lane to a pseudo-random slot of a 1024-slot table, the
//                   z-slot keys of a 64 x 32-cell tile
//   global_min_u64  64-bit global atomicMin without return, pseudo-random
//                   slots of a 64 MiB table (8M slots, a 4K frame's keys)
//   global_add_f64  binary64 global atomicAdd without return, pseudo-random
//                   entries of a 1.2 MB table (the shading gradients of a
//                   50,000-vertex mesh)
//   lds_add_f32     binary32 LDS atomicAdd without return (ds_add_f32),
//                   pseudo-random slots of a 1024-slot table
//   lds_add_u64     64-bit integer LDS atomicAdd without return (ds_add_u64),
//                   the same pattern (the shading backward's fixed-point
//                   vertex sums)
//
// Prints one JSON object: operations per second for each (best of 5 runs),
// and the grid.  Every address is masked into its table, so no access leaves
// its buffer.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

constexpr int BLOCK = 256;
constexpr int LDS_SLOTS = 1024;
constexpr uint32_t GLOBAL_SLOTS = 1u << 23;     // 64 MiB of u64
constexpr uint32_t ADD_SLOTS = 150000;          // ~50k vertices x 3

__device__ __forceinline__ uint32_t step(uint32_t h) { return h * 1664525u + 1013904223u; }

__global__ void __launch_bounds__(BLOCK) k_lds_min(uint64_t* out, int iters)
{
    __shared__ uint64_t keys[LDS_SLOTS];
    for (int i = threadIdx.x; i < LDS_SLOTS; i += BLOCK) keys[i] = ~0ull;
    __syncthreads();
    uint32_t h = (blockIdx.x * BLOCK + threadIdx.x) * 2654435761u + 12345u;
    for (int k = 0; k < iters; ++k) {
        h = step(h);
        atomicMin((unsigned long long*)&keys[h >> 22], ((unsigned long long)h << 16) | (unsigned)k);
    }
    __syncthreads();
    // fold the table into one word per block so the atomics are not dead
    if (threadIdx.x == 0) {
        uint64_t m = ~0ull;
        for (int i = 0; i < LDS_SLOTS; ++i) m = m < keys[i] ? m : keys[i];
        out[blockIdx.x] = m;
    }
}

template <class T>
__global__ void __launch_bounds__(BLOCK) k_lds_add(uint64_t* out, int iters)
{
    __shared__ T sums[LDS_SLOTS];
    for (int i = threadIdx.x; i < LDS_SLOTS; i += BLOCK) sums[i] = 0;
    __syncthreads();
    uint32_t h = (blockIdx.x * BLOCK + threadIdx.x) * 2654435761u + 4242u;
    for (int k = 0; k < iters; ++k) {
        h = step(h);
        atomicAdd(&sums[h >> 22], (T)(h & 7u));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        T m = 0;
        for (int i = 0; i < LDS_SLOTS; ++i) m += sums[i];
        out[blockIdx.x] = (uint64_t)m;
    }
}

__global__ void __launch_bounds__(BLOCK) k_global_min(uint64_t* table, int iters)
{
    uint32_t h = (blockIdx.x * BLOCK + threadIdx.x) * 2654435761u + 777u;
    for (int k = 0; k < iters; ++k) {
        h = step(h);
        atomicMin((unsigned long long*)&table[h & (GLOBAL_SLOTS - 1)], ((unsigned long long)h << 16) | (unsigned)k);
    }
}

__global__ void __launch_bounds__(BLOCK) k_global_add(double* table, int iters)
{
    uint32_t h = (blockIdx.x * BLOCK + threadIdx.x) * 2654435761u + 99u;
    for (int k = 0; k < iters; ++k) {
        h = step(h);
        atomicAdd(&table[(h >> 8) % ADD_SLOTS], 1.0);
    }
}

}  // namespace

#define CHECK(x)                                                                          \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            fprintf(stderr, "atomic_peak: %s: %s\n", #x, hipGetErrorString(e_));          \
            return 1;                                                                     \
        }                                                                                 \
    } while (0)

int main()
{
    hipDeviceProp_t p;
    CHECK(hipGetDeviceProperties(&p, 0));
    const int grid = p.multiProcessorCount * 16;
    uint64_t *out = nullptr, *table = nullptr;
    double* acc = nullptr;
    CHECK(hipMalloc(&out, sizeof(uint64_t) * grid));
    CHECK(hipMalloc(&table, sizeof(uint64_t) * GLOBAL_SLOTS));
    CHECK(hipMalloc(&acc, sizeof(double) * ADD_SLOTS));
    CHECK(hipMemset(table, 0xFF, sizeof(uint64_t) * GLOBAL_SLOTS));
    CHECK(hipMemset(acc, 0, sizeof(double) * ADD_SLOTS));
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    double rate[5];
    const int iters[5] = {4096, 256, 64, 512, 4096};
    for (int which = 0; which < 5; ++which) {
        float best = 1e30f;
        for (int rep = 0; rep < 6; ++rep) {   // the first run warms up
            CHECK(hipEventRecord(a, 0));
            if (which == 0) hipLaunchKernelGGL(k_lds_min, dim3(grid), dim3(BLOCK), 0, 0, out, iters[0]);
            if (which == 1) hipLaunchKernelGGL(k_global_min, dim3(grid), dim3(BLOCK), 0, 0, table, iters[1]);
            if (which == 2) hipLaunchKernelGGL(k_global_add, dim3(grid), dim3(BLOCK), 0, 0, acc, iters[2]);
            if (which == 3) hipLaunchKernelGGL(k_lds_add<float>, dim3(grid), dim3(BLOCK), 0, 0, out, iters[3]);
            if (which == 4)
                hipLaunchKernelGGL(k_lds_add<unsigned long long>, dim3(grid), dim3(BLOCK), 0, 0, out, iters[4]);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(b, 0));
            CHECK(hipEventSynchronize(b));
            float ms = 0.0f;
            CHECK(hipEventElapsedTime(&ms, a, b));
            if (rep > 0 && ms < best) best = ms;
        }
        rate[which] = (double)grid * BLOCK * iters[which] / (best * 1e-3);
    }
    printf("{\"lds_min_u64_per_s\": %.4e, \"global_min_u64_per_s\": %.4e, \"global_add_f64_per_s\": %.4e, "
           "\"lds_add_f32_per_s\": %.4e, \"lds_add_u64_per_s\": %.4e, "
           "\"grid\": %d, \"block\": %d, \"cus\": %d}\n",
           rate[0], rate[1], rate[2], rate[3], rate[4], grid, BLOCK, p.multiProcessorCount);
    CHECK(hipFree(out));
    CHECK(hipFree(table));
    CHECK(hipFree(acc));
    return 0;
}
