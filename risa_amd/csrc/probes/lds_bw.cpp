// LDS read bandwidth of gfx950, the ceiling the backprojectors' LDS gathers
// are measured against.  Every workgroup fills 32 KB of LDS, then each lane
// reads 16/8/4-byte entries (consecutive lanes at consecutive addresses, the
// index masked into the buffer: no bank conflicts, all accesses in bounds)
// and folds them with xor so the reads stay live; one warm-up launch, then
// one timed launch of 4 workgroups per CU.  Prints chip-wide TB/s per width.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

template <int W>   // dwords per read: 4, 2, 1
__global__ void __launch_bounds__(256) lds_read(unsigned* out, int iters)
{
    __shared__ unsigned buf[8192];   // 32 KB
    const int tid = threadIdx.x;
    for (int i = tid; i < 8192; i += 256) buf[i] = i * 2654435761u;
    __syncthreads();
    unsigned a0 = 0, a1 = 0;
    constexpr int n = 8192 / W;   // entries
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = (tid + (it * 4 + u) * 64) & (n - 1);
            if (W == 4) {
                const uint4 v = reinterpret_cast<const uint4*>(buf)[e];
                a0 ^= v.x ^ v.y;
                a1 ^= v.z ^ v.w;
            } else if (W == 2) {
                const uint2 v = reinterpret_cast<const uint2*>(buf)[e];
                a0 ^= v.x;
                a1 ^= v.y;
            } else {
                a0 ^= buf[e];
            }
        }
    }
    out[blockIdx.x * 256 + tid] = a0 ^ a1;
}

static void check(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
}

template <int W>
static void run(int wgs, int iters, unsigned* out)
{
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "hipEventCreate");
    check(hipEventCreate(&e1), "hipEventCreate");
    hipLaunchKernelGGL(lds_read<W>, dim3(wgs), dim3(256), 0, 0, out, 16);
    check(hipDeviceSynchronize(), "warm-up");
    check(hipEventRecord(e0, 0), "hipEventRecord");
    hipLaunchKernelGGL(lds_read<W>, dim3(wgs), dim3(256), 0, 0, out, iters);
    check(hipEventRecord(e1, 0), "hipEventRecord");
    check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0;
    check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    const double bytes = (double)wgs * 256 * iters * 4 * W * 4;
    std::printf("{\"read_bytes\": %d, \"ms\": %.3f, \"tb_per_s\": %.1f}\n", 4 * W, ms, bytes / ms / 1e9);
    check(hipEventDestroy(e0), "hipEventDestroy");
    check(hipEventDestroy(e1), "hipEventDestroy");
}

int main()
{
    hipDeviceProp_t p;
    check(hipGetDeviceProperties(&p, 0), "hipGetDeviceProperties");
    std::printf("{\"device\": \"%s\", \"cus\": %d, \"clock_khz\": %d}\n", p.gcnArchName, p.multiProcessorCount,
                p.clockRate);
    const int wgs = p.multiProcessorCount * 4;
    unsigned* out = nullptr;
    check(hipMalloc(&out, (size_t)wgs * 256 * sizeof(unsigned)), "hipMalloc");
    run<4>(wgs, 8192, out);
    run<2>(wgs, 8192, out);
    run<1>(wgs, 8192, out);
    check(hipFree(out), "hipFree");
    return 0;
}
