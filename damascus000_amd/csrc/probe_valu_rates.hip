// VALU issue-rate probe for gfx950: 32-bit mul_lo / mul_hi, mad_u64_u32, add, fp64 fma, rndne.
// Eight independent dependency chains per lane; reports lane-ops per second against
// 256 CU x 64 lanes x 2.4 GHz. Build: make -C damascus000_amd/csrc probe; run: damascus000_amd/probe_valu_rates
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>

#define ITERS 4096
#define CHAINS 8

__global__ void k_mullo(uint32_t* out, uint32_t seed) {
    uint32_t x[CHAINS];
    uint32_t y = seed | 0x80000001u;
    for (int c = 0; c < CHAINS; ++c) x[c] = threadIdx.x * 2654435761u + c * 40503u + seed;
    for (int i = 0; i < ITERS; ++i) {
#pragma unroll
        for (int c = 0; c < CHAINS; ++c) x[c] = x[c] * (x[c] | y);
    }
    uint32_t s = 0;
    for (int c = 0; c < CHAINS; ++c) s ^= x[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void k_mulhi(uint32_t* out, uint32_t seed) {
    uint32_t x[CHAINS];
    uint32_t y = seed | 0x80000001u;
    for (int c = 0; c < CHAINS; ++c) x[c] = threadIdx.x * 2654435761u + c * 40503u + seed;
    for (int i = 0; i < ITERS; ++i) {
#pragma unroll
        for (int c = 0; c < CHAINS; ++c) x[c] = __umulhi(x[c], y) ^ y;
    }
    uint32_t s = 0;
    for (int c = 0; c < CHAINS; ++c) s ^= x[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void k_mad64(uint32_t* out, uint32_t seed) {
    uint64_t x[CHAINS];
    uint32_t y = seed | 0x80000001u;
    for (int c = 0; c < CHAINS; ++c) x[c] = threadIdx.x * 2654435761u + c * 40503u + seed;
    for (int i = 0; i < ITERS; ++i) {
#pragma unroll
        for (int c = 0; c < CHAINS; ++c) x[c] = (uint64_t)(uint32_t)x[c] * y + (x[c] >> 32);
    }
    uint32_t s = 0;
    for (int c = 0; c < CHAINS; ++c) s ^= (uint32_t)x[c] ^ (uint32_t)(x[c] >> 32);
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void k_add(uint32_t* out, uint32_t seed) {
    uint32_t x[CHAINS];
    uint32_t y = seed | 0x80000001u;
    for (int c = 0; c < CHAINS; ++c) x[c] = threadIdx.x * 2654435761u + c * 40503u + seed;
    for (int i = 0; i < ITERS; ++i) {
#pragma unroll
        for (int c = 0; c < CHAINS; ++c) x[c] = (x[c] + y) ^ c;
    }
    uint32_t s = 0;
    for (int c = 0; c < CHAINS; ++c) s ^= x[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void k_fma64(uint32_t* out, uint32_t seed) {
    double x[CHAINS];
    double y = 1.0000001 + seed * 1e-12;
    for (int c = 0; c < CHAINS; ++c) x[c] = threadIdx.x * 0.001 + c;
    for (int i = 0; i < ITERS; ++i) {
#pragma unroll
        for (int c = 0; c < CHAINS; ++c) x[c] = __builtin_fma(x[c], y, -0.5);
    }
    double s = 0;
    for (int c = 0; c < CHAINS; ++c) s += x[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = (uint32_t)(int64_t)s;
}

__global__ void k_rint64(uint32_t* out, uint32_t seed) {
    double x[CHAINS];
    double y = 1.0000001 + seed * 1e-12;
    for (int c = 0; c < CHAINS; ++c) x[c] = threadIdx.x * 0.001 + c;
    for (int i = 0; i < ITERS; ++i) {
#pragma unroll
        for (int c = 0; c < CHAINS; ++c) x[c] = __builtin_rint(x[c] * y) + 0.25;
    }
    double s = 0;
    for (int c = 0; c < CHAINS; ++c) s += x[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = (uint32_t)(int64_t)s;
}

typedef void (*kfn)(uint32_t*, uint32_t);

static void run(const char* name, kfn f, int ops_per_iter_chain, uint32_t* d) {
    const int blocks = 256 * 8, threads = 256;
    hipEvent_t a, b;
    hipEventCreate(&a);
    hipEventCreate(&b);
    hipLaunchKernelGGL(f, dim3(blocks), dim3(threads), 0, 0, d, 12345u);
    hipDeviceSynchronize();
    hipEventRecord(a);
    for (int r = 0; r < 5; ++r) hipLaunchKernelGGL(f, dim3(blocks), dim3(threads), 0, 0, d, 12345u + r);
    hipEventRecord(b);
    hipEventSynchronize(b);
    float ms = 0;
    hipEventElapsedTime(&ms, a, b);
    double lane_ops = 5.0 * blocks * threads * (double)ITERS * CHAINS * ops_per_iter_chain;
    printf("%-8s %.3f ms  %.2f T lane-ops/s  (%.1f%% of 39.32T)\n", name, ms, lane_ops / (ms * 1e-3) / 1e12,
           100.0 * lane_ops / (ms * 1e-3) / 39.3216e12);
    fflush(stdout);
}

int main() {
    uint32_t* d;
    hipMalloc(&d, 256 * 8 * 256 * 4);
    hipDeviceProp_t prop;
    hipGetDeviceProperties(&prop, 0);
    printf("device %s CUs %d clock %d kHz lds/block %zu\n", prop.gcnArchName, prop.multiProcessorCount,
           prop.clockRate, prop.sharedMemPerBlock);
    run("add", k_add, 2, d);
    run("mullo+or", k_mullo, 2, d);
    run("mulhi", k_mulhi, 2, d);
    run("mad64", k_mad64, 2, d);
    run("fma64", k_fma64, 1, d);
    run("rint64", k_rint64, 3, d);
    hipFree(d);
    return 0;
}
