// Test probe (built by __graft_entry__.build(), run by tests/test_gpu_parity.py).
// Exhaustive check: v_cvt_rpi_i32_f32(x) == int(double(x) + 0.5) for every float x in
// [-0.5, 2^24] (bit patterns from -0.5 down to -0 and from +0 up to 2^24).
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
#include <cstring>
__global__ void k_check(uint32_t first, uint32_t count, unsigned long long* bad, uint32_t* example) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= count) return;
    const uint32_t bits = first + i;
    const float x = __builtin_bit_cast(float, bits);
    int r;
    asm volatile("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    const int ref = int(double(x) + 0.5);
    if (r != ref) {
        atomicAdd(bad, 1ull);
        atomicMin(example, bits);
    }
}
int main() {
    unsigned long long* bad; uint32_t* ex;
    if (hipMalloc(&bad, 8) != hipSuccess || hipMalloc(&ex, 4) != hipSuccess) return 2;
    // ranges of bit patterns: positive [0x00000000, bits(2^24)], negative [0x80000000, bits(-0.5)]
    float lim = 16777216.0f, mhalf = -0.5f;
    uint32_t lim_bits, mhalf_bits; std::memcpy(&lim_bits, &lim, 4); std::memcpy(&mhalf_bits, &mhalf, 4);
    struct R { uint32_t a, b; } ranges[2] = {{0u, lim_bits}, {0x80000000u, mhalf_bits}};
    unsigned long long total_bad = 0, total = 0;
    for (auto rg : ranges) {
        unsigned long long zero = 0; uint32_t big = 0xffffffffu;
        hipMemcpy(bad, &zero, 8, hipMemcpyHostToDevice); hipMemcpy(ex, &big, 4, hipMemcpyHostToDevice);
        for (uint64_t s = rg.a; s <= rg.b; s += (1u << 28)) {
            const uint32_t count = uint32_t(std::min<uint64_t>(1u << 28, uint64_t(rg.b) - s + 1));
            k_check<<<(count + 255) / 256, 256>>>(uint32_t(s), count, bad, ex);
            total += count;
        }
        if (hipDeviceSynchronize() != hipSuccess) { printf("sync failed\n"); return 3; }
        unsigned long long b; uint32_t e;
        hipMemcpy(&b, bad, 8, hipMemcpyDeviceToHost); hipMemcpy(&e, ex, 4, hipMemcpyDeviceToHost);
        float ef; std::memcpy(&ef, &e, 4);
        printf("range %08x..%08x: %llu mismatches%s", rg.a, rg.b, b, b ? "" : "\n");
        if (b) printf(" (smallest bits %08x = %.9g)\n", e, ef);
        total_bad += b;
    }
    printf("checked %llu floats, %llu mismatches\n", total, total_bad);
    return total_bad ? 1 : 0;
}
