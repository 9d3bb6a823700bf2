// Measured memory rates of one GPU, the ceilings the grid encodings run into:
//   - random gathers: every thread issues K loads of W bytes at hashed
//     (uniformly spread) indices into a table of S bytes; reported as lane
//     loads per clock per CU and per second, for tables from L1-sized to
//     beyond the L2 and the Infinity Cache;
//   - streaming: a read-only sum and a copy over 1 GiB buffers (HBM).
// Build: hipcc --offload-arch=gfx950 -O3 gather_rate.hip -o gather_rate
// (guiding-df_amd/build.py builds it beside the library).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                             \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            std::exit(1);                                                                 \
        }                                                                                 \
    } while (0)

// integer finaliser (lowbias32): uniformly spread indices without a table
__device__ __forceinline__ uint32_t spread(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// K loads of W bytes per thread at indices masked into `entries` 16-byte slots.
template <int K, int W>
__global__ __launch_bounds__(256) void gather_kernel(const uint4* __restrict__ tab, uint32_t slot_mask,
                                                     float* __restrict__ out, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint32_t idx[K];
#pragma unroll
    for (int k = 0; k < K; ++k) idx[k] = spread((uint32_t)t * 16u + (uint32_t)k) & slot_mask;
    float acc = 0.0f;
    if constexpr (W == 16) {
        uint4 v[K];
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = tab[idx[k]];
#pragma unroll
        for (int k = 0; k < K; ++k) acc += __uint_as_float(v[k].x) + __uint_as_float(v[k].w);
    } else if constexpr (W == 8) {
        const uint2* t2 = (const uint2*)tab;
        uint2 v[K];
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = t2[(size_t)idx[k] * 2];
#pragma unroll
        for (int k = 0; k < K; ++k) acc += __uint_as_float(v[k].x) + __uint_as_float(v[k].y);
    } else {
        const uint32_t* t1 = (const uint32_t*)tab;
        uint32_t v[K];
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = t1[(size_t)idx[k] * 4];
#pragma unroll
        for (int k = 0; k < K; ++k) acc += __uint_as_float(v[k]);
    }
    out[t] = acc;
}

__global__ __launch_bounds__(256) void read_kernel(const float4* __restrict__ src, int64_t n4, float* __restrict__ out) {
    float acc = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x) {
        const float4 v = src[i];
        acc += v.x + v.y + v.z + v.w;
    }
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

__global__ __launch_bounds__(256) void copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n4) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

struct Timer {
    hipEvent_t a, b;
    Timer() { CK(hipEventCreate(&a)); CK(hipEventCreate(&b)); }
    ~Timer() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    void start() { CK(hipEventRecord(a)); }
    float stop_ms() {
        CK(hipEventRecord(b));
        CK(hipEventSynchronize(b));
        float ms = 0.0f;
        CK(hipEventElapsedTime(&ms, a, b));
        return ms;
    }
};

static int g_cus = 0;
static double g_ghz = 0.0;

template <int K, int W>
static void gather(const uint4* tab, size_t bytes, float* out, int64_t n) {
    const uint32_t slots = (uint32_t)(bytes / 16);  // a power of two
    const unsigned blocks = (unsigned)((n + 255) / 256);
    for (int i = 0; i < 3; ++i)
        hipLaunchKernelGGL((gather_kernel<K, W>), dim3(blocks), dim3(256), 0, 0, tab, slots - 1, out, n);
    CK(hipGetLastError());
    Timer t;
    const int reps = 10;
    t.start();
    for (int i = 0; i < reps; ++i)
        hipLaunchKernelGGL((gather_kernel<K, W>), dim3(blocks), dim3(256), 0, 0, tab, slots - 1, out, n);
    const double us = t.stop_ms() * 1000.0 / reps;
    const double loads = (double)n * K;
    std::printf("%10zu KB  %2d B  %2d  %9.1f  %8.1f  %6.3f\n", bytes / 1024, W, K, us, loads / us / 1e3,
                loads / (us * 1e-6 * g_ghz * 1e9) / g_cus);
    std::fflush(stdout);
}

int main() {
    int khz = 0;
    CK(hipDeviceGetAttribute(&g_cus, hipDeviceAttributeMultiprocessorCount, 0));
    CK(hipDeviceGetAttribute(&khz, hipDeviceAttributeClockRate, 0));
    g_ghz = khz / 1e6;
    hipDeviceProp_t prop;
    CK(hipGetDeviceProperties(&prop, 0));
    std::printf("device %s, %d CUs, %.2f GHz (peak clock attribute), L2 %d KB\n", prop.gcnArchName, g_cus, g_ghz,
                prop.l2CacheSize / 1024);

    const int64_t n = (int64_t)1 << 24;  // threads per gather launch
    const size_t table_max = (size_t)64 << 20;
    uint4* tab = nullptr;
    float* out = nullptr;
    CK(hipMalloc(&tab, table_max));
    CK(hipMalloc(&out, n * sizeof(float)));
    CK(hipMemset(tab, 0, table_max));
    std::printf("\nrandom gathers (%lld threads, K loads each)\n", (long long)n);
    std::printf("%13s  %4s  %2s  %9s  %8s  %6s\n", "table", "W", "K", "us", "G ld/s", "ld/clk/CU");
    for (size_t kb : {32, 128, 1024, 2048, 4096, 8192, 16384, 65536}) {
        const size_t bytes = kb * 1024;
        gather<8, 4>(tab, bytes, out, n);
        gather<8, 8>(tab, bytes, out, n);
        gather<8, 16>(tab, bytes, out, n);
    }
    CK(hipFree(tab));

    // streaming over 1 GiB
    const size_t sb = (size_t)1 << 30;
    const int64_t n4 = (int64_t)(sb / 16);
    float4 *src = nullptr, *dst = nullptr;
    CK(hipMalloc(&src, sb));
    CK(hipMalloc(&dst, sb));
    CK(hipMemset(src, 0, sb));
    const unsigned sblocks = (unsigned)(g_cus * 16);
    float* part = nullptr;
    CK(hipMalloc(&part, (size_t)sblocks * 256 * sizeof(float)));
    Timer t;
    for (int i = 0; i < 2; ++i) hipLaunchKernelGGL(read_kernel, dim3(sblocks), dim3(256), 0, 0, src, n4, part);
    t.start();
    for (int i = 0; i < 5; ++i) hipLaunchKernelGGL(read_kernel, dim3(sblocks), dim3(256), 0, 0, src, n4, part);
    const double rd_ms = t.stop_ms() / 5;
    for (int i = 0; i < 2; ++i) hipLaunchKernelGGL(copy_kernel, dim3(sblocks), dim3(256), 0, 0, src, dst, n4);
    t.start();
    for (int i = 0; i < 5; ++i) hipLaunchKernelGGL(copy_kernel, dim3(sblocks), dim3(256), 0, 0, src, dst, n4);
    const double cp_ms = t.stop_ms() / 5;
    CK(hipGetLastError());
    std::printf("\nstreaming, 1 GiB: read %.0f GB/s, copy %.0f GB/s (read + write bytes)\n", sb / (rd_ms * 1e6),
                2.0 * sb / (cp_ms * 1e6));
    CK(hipFree(src));
    CK(hipFree(dst));
    CK(hipFree(part));
    CK(hipFree(out));
    return 0;
}
