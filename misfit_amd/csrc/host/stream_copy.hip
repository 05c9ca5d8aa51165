// stream_copy -- STREAM-style device copy of a known size on one GPU.
//
// Reports the copy bandwidth the HBM actually sustains (read + write bytes /
// time) and serves as the calibration kernel for hardware byte counters: a
// counter run over this program reads exactly `bytes` per dispatch and
// writes exactly `bytes`, so counter/bytes is the scale factor to apply to
// the same counter on any other kernel.
//   stream_copy [MiB per buffer, default 2048] [repetitions, default 20]
// Prints one JSON line.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

#define CHK(x)                                                                          \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) {                                                         \
            fprintf(stderr, "stream_copy: %s failed: %s\n", #x, hipGetErrorString(e_)); \
            return 1;                                                                   \
        }                                                                               \
    } while (0)

__global__ __launch_bounds__(256) void k_copy(const int4* __restrict__ src, int4* __restrict__ dst, long long count) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < count; i += stride) dst[i] = src[i];
}

int main(int argc, char** argv) {
    long long mib = argc > 1 ? atoll(argv[1]) : 2048;
    int reps = argc > 2 ? atoi(argv[2]) : 20;
    if (mib <= 0 || reps <= 0) {
        fprintf(stderr, "usage: stream_copy [MiB per buffer] [repetitions]\n");
        return 1;
    }
    size_t bytes = (size_t)mib << 20;
    long long count = (long long)(bytes / sizeof(int4));
    int4 *a = nullptr, *b = nullptr;
    CHK(hipMalloc(&a, bytes));
    CHK(hipMalloc(&b, bytes));
    CHK(hipMemset(a, 1, bytes));
    CHK(hipMemset(b, 0, bytes));
    hipDeviceProp_t prop;
    CHK(hipGetDeviceProperties(&prop, 0));
    int blocks = prop.multiProcessorCount * 8;
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    k_copy<<<blocks, 256>>>(a, b, count);
    CHK(hipGetLastError());
    CHK(hipDeviceSynchronize());
    CHK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; r++) k_copy<<<blocks, 256>>>(a, b, count);
    CHK(hipEventRecord(e1, 0));
    CHK(hipEventSynchronize(e1));
    CHK(hipGetLastError());
    float ms = 0.f;
    CHK(hipEventElapsedTime(&ms, e0, e1));
    double per = (double)ms / reps;
    double gbs = 2.0 * (double)bytes / (per * 1e-3) / 1e9;
    printf("{\"kernel\": \"k_copy\", \"bytes_read_per_dispatch\": %zu, \"bytes_written_per_dispatch\": %zu, "
           "\"ms_per_dispatch\": %.4f, \"copy_gbs\": %.1f, \"reps\": %d}\n",
           bytes, bytes, per, gbs, reps + 1);
    CHK(hipFree(a));
    CHK(hipFree(b));
    return 0;
}
