// A plain streaming kernel of known traffic, used to calibrate the HBM counters (FETCH_SIZE,
// WRITE_SIZE) and to measure the bandwidth a two-reads-one-write stream reaches on this device.
// No reference counterpart; nothing on the stencil path calls it.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace pm {

// c[i] = a[i] + b[i]: 16 bytes read and 8 written per element, the heat sweep's mix.
__global__ __launch_bounds__(256) void stream_add(const double *__restrict__ a,
                                                  const double *__restrict__ b,
                                                  double *__restrict__ c, int64_t n) {
    for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        c[i] = a[i] + b[i];
}

}  // namespace pm

using namespace pm;

extern "C" int pm_stream_copy(size_t n, int reps, double *ms_per_launch) {
    PM_CHECK(require_device());
    if (n == 0 || reps <= 0) return PM_ERR_ARGUMENT;
    DevBuf a, b, c;
    PM_CHECK(a.alloc(n * sizeof(double)));
    PM_CHECK(b.alloc(n * sizeof(double)));
    PM_CHECK(c.alloc(n * sizeof(double)));
    PM_HIP(hipMemset(a.d(), 0, n * sizeof(double)));
    PM_HIP(hipMemset(b.d(), 0, n * sizeof(double)));
    hipEvent_t e0, e1;
    PM_HIP(hipEventCreate(&e0));
    PM_HIP(hipEventCreate(&e1));
    const int grid = 65536;
    stream_add<<<grid, 256>>>(a.d(), b.d(), c.d(), (int64_t)n);  // warm
    PM_HIP(hipEventRecord(e0));
    for (int r = 0; r < reps; r++) stream_add<<<grid, 256>>>(a.d(), b.d(), c.d(), (int64_t)n);
    PM_HIP(hipEventRecord(e1));
    PM_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    PM_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_per_launch) *ms_per_launch = ms / reps;
    return PM_OK;
}
