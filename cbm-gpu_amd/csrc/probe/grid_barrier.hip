// Measures what one level of a tree pass would cost as a grid barrier inside
// a persistent cooperative kernel versus as a dependent kernel launch on the
// same stream, for grids of 1, 2, 4 and 8 workgroups per CU; the barrier both
// as cooperative groups' grid sync and as a hand-rolled atomic barrier.  The
// CBM product runs one launch per tree level because of this measurement.
//   make -C cbm-gpu_amd/csrc probe && cbm-gpu_amd/csrc/probe/grid_barrier
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>
#include <cstdio>
namespace cg = cooperative_groups;

__global__ __launch_bounds__(256) void k_sync(int L, float *buf) {
    cg::grid_group g = cg::this_grid();
    for (int l = 0; l < L; ++l) {
        if (threadIdx.x == 0) buf[blockIdx.x] += 1.0f;
        g.sync();
    }
}

// A hand-rolled barrier: one arrival counter and a generation word in device
// memory, agent-scope atomics; the last workgroup to arrive resets the counter
// and advances the generation, the others poll it.  The poll gives up after
// `spin_limit` rounds and sets *err, so the kernel always ends (the grid is
// co-resident under the cooperative launch, so it never should).
__global__ __launch_bounds__(256) void k_atomic_sync(int L, float *buf, unsigned *count, unsigned *gen,
                                                     unsigned spin_limit, int *err) {
    __shared__ unsigned s_gen;
    if (threadIdx.x == 0) s_gen = __hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    for (int l = 0; l < L; ++l) {
        if (threadIdx.x == 0) buf[blockIdx.x] += 1.0f;
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned g = s_gen;
            const unsigned arrived = __hip_atomic_fetch_add(count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (arrived == gridDim.x - 1) {
                __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                unsigned spins = 0;
                while (__hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
                    if (++spins > spin_limit) {
                        atomicOr(err, 1);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            s_gen = g + 1;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_one(float *buf) {
    if (threadIdx.x == 0) buf[blockIdx.x] += 1.0f;
}

#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("%s: %s\n", #x, hipGetErrorString(e)); return 1; } } while (0)

int main() {
    int dev = 0, cus = 0, nb = 0;
    CK(hipGetDevice(&dev));
    CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sync, 256, 0));
    printf("CUs %d, max co-resident blocks per CU %d\n", cus, nb);
    float *buf;
    CK(hipMalloc(&buf, sizeof(float) * 65536));
    CK(hipMemset(buf, 0, sizeof(float) * 65536));
    unsigned *words;  // count, generation
    int *err;
    CK(hipMalloc(&words, 2 * sizeof(unsigned)));
    CK(hipMalloc(&err, sizeof(int)));
    CK(hipMemset(words, 0, 2 * sizeof(unsigned)));
    CK(hipMemset(err, 0, sizeof(int)));
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    const int L = 200;
    for (int grid : {cus, 2 * cus, 4 * cus, nb * cus}) {
        if (grid > nb * cus || grid > 65536) continue;
        int Lv = L;
        void *args[] = {&Lv, &buf};
        CK(hipLaunchCooperativeKernel((void *)k_sync, dim3(grid), dim3(256), args, 0, 0));
        CK(hipDeviceSynchronize());
        CK(hipEventRecord(a, 0));
        CK(hipLaunchCooperativeKernel((void *)k_sync, dim3(grid), dim3(256), args, 0, 0));
        CK(hipEventRecord(b, 0));
        CK(hipEventSynchronize(b));
        float ms;
        CK(hipEventElapsedTime(&ms, a, b));
        printf("grid %6d: %8.2f us per grid barrier\n", grid, 1e3 * ms / L);
        {
            unsigned *count = words, *gen = words + 1;
            unsigned spin_limit = 1u << 20;
            void *aargs[] = {&Lv, &buf, &count, &gen, &spin_limit, &err};
            CK(hipLaunchCooperativeKernel((void *)k_atomic_sync, dim3(grid), dim3(256), aargs, 0, 0));
            CK(hipDeviceSynchronize());
            CK(hipEventRecord(a, 0));
            CK(hipLaunchCooperativeKernel((void *)k_atomic_sync, dim3(grid), dim3(256), aargs, 0, 0));
            CK(hipEventRecord(b, 0));
            CK(hipEventSynchronize(b));
            CK(hipEventElapsedTime(&ms, a, b));
            int e = 0;
            CK(hipMemcpy(&e, err, sizeof(int), hipMemcpyDeviceToHost));
            printf("grid %6d: %8.2f us per hand-rolled atomic barrier%s\n", grid, 1e3 * ms / L,
                   e ? " (a poll gave up: timing invalid)" : "");
        }
        for (int i = 0; i < 3; ++i) k_one<<<grid, 256>>>(buf);
        CK(hipDeviceSynchronize());
        CK(hipEventRecord(a, 0));
        for (int i = 0; i < L; ++i) k_one<<<grid, 256>>>(buf);
        CK(hipEventRecord(b, 0));
        CK(hipEventSynchronize(b));
        CK(hipEventElapsedTime(&ms, a, b));
        printf("grid %6d: %8.2f us per dependent launch\n", grid, 1e3 * ms / L);
    }
    CK(hipFree(buf));
    CK(hipFree(words));
    CK(hipFree(err));
    return 0;
}
