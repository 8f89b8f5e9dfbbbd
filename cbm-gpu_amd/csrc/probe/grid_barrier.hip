// Measures what one level of a tree pass would cost as a grid barrier inside
// a persistent cooperative kernel versus as a dependent kernel launch on the
// same stream, for grids of 1, 2, 4 and 8 workgroups per CU.  The CBM product
// runs one launch per tree level because of this measurement.
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
    return 0;
}
