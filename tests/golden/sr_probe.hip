// Writes the fixture of gfx950's stochastic-rounding conversions
// (v_cvt_sr_f16_f32, v_cvt_sr_bf16_f32): one (x, r) pair per thread, a
// bounds-checked grid over the pairs, the two 16-bit results copied back.
// The pairs: 160 random (x, r) for every float exponent of x from -45 to
// 17; 40 values of x against the structured words 0, ~0, 1 << k,
// (1 << k) - 1 and ~((1 << k) - 1); zeros, binary16's limits and the
// powers of two around its subnormal range against 16 words each; inf and
// NaN.  Build and run with make_sr_fixture.py.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <vector>

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf162_t __attribute__((ext_vector_type(2)));

__global__ void convert(const uint32_t* x, const uint32_t* r, uint16_t* h, uint16_t* b, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float xf = __builtin_bit_cast(float, x[i]);
    h[i] = __builtin_bit_cast(uint16_t, __builtin_amdgcn_cvt_sr_f16_f32(half2_t{}, xf, r[i], false)[0]);
    b[i] = __builtin_bit_cast(uint16_t, __builtin_amdgcn_cvt_sr_bf16_f32(bf162_t{}, xf, r[i], false)[0]);
}

static uint64_t state = 0x1234567887654321ull;
static uint64_t splitmix64() {
    uint64_t z = (state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static uint32_t random_float(int e) {   // random sign and significand, exponent e
    return ((uint32_t)(splitmix64() & 1) << 31) | ((uint32_t)(e + 127) << 23) | (uint32_t)(splitmix64() & 0x7fffff);
}

#define CHECK(c) do { hipError_t err_ = (c); if (err_ != hipSuccess) { \
    fprintf(stderr, "%s: %s\n", #c, hipGetErrorString(err_)); return 1; } } while (0)

int main(int argc, char** argv) {
    if (argc != 2) { fprintf(stderr, "usage: sr_probe OUT\n"); return 2; }
    std::vector<uint32_t> X, R;
    auto add = [&](uint32_t x, uint32_t r) { X.push_back(x); R.push_back(r); };
    for (int e = -45; e <= 17; e++)
        for (int j = 0; j < 160; j++) add(random_float(e), (uint32_t)splitmix64());
    for (int j = 0; j < 40; j++) {
        const uint32_t x = random_float(-30 + (int)(splitmix64() % 47));
        add(x, 0u);
        add(x, ~0u);
        for (int k = 0; k < 32; k++) {
            add(x, 1u << k);
            add(x, (1u << k) - 1u);
            add(x, ~((1u << k) - 1u));
        }
    }
    const float special[] = {0.f, 65504.f, 65519.f, 65520.f, 65535.f, 65536.f, 1e10f, 6.103515625e-05f,
                             5.9604644775390625e-08f, 2.98023223876953125e-08f, 1.4901161193847656e-08f,
                             7.275957614183426e-12f, 3e-8f, 4e-8f, 1e-30f, 3.0e38f, 1.0f, 0.5f};
    for (float f : special)
        for (int sign = 0; sign < 2; sign++)
            for (int j = 0; j < 16; j++)
                add(__builtin_bit_cast(uint32_t, sign ? -f : f), j < 8 ? (uint32_t)splitmix64() : (uint32_t)j << 28);
    for (uint32_t u : {0x7f800000u, 0xff800000u, 0x7fc00000u, 0x7fa00000u})
        for (int j = 0; j < 4; j++) add(u, (uint32_t)splitmix64());
    const int n = (int)X.size();
    uint32_t *dx, *dr;
    uint16_t *dh, *db;
    CHECK(hipMalloc(&dx, n * sizeof(uint32_t)));
    CHECK(hipMalloc(&dr, n * sizeof(uint32_t)));
    CHECK(hipMalloc(&dh, n * sizeof(uint16_t)));
    CHECK(hipMalloc(&db, n * sizeof(uint16_t)));
    CHECK(hipMemcpy(dx, X.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    CHECK(hipMemcpy(dr, R.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(convert, dim3((n + 255) / 256), dim3(256), 0, 0, dx, dr, dh, db, n);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    std::vector<uint16_t> H(n), B(n);
    CHECK(hipMemcpy(H.data(), dh, n * sizeof(uint16_t), hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(B.data(), db, n * sizeof(uint16_t), hipMemcpyDeviceToHost));
    FILE* f = fopen(argv[1], "wb");
    if (!f) return 1;
    fwrite(&n, sizeof(int), 1, f);
    fwrite(X.data(), sizeof(uint32_t), n, f);
    fwrite(R.data(), sizeof(uint32_t), n, f);
    fwrite(H.data(), sizeof(uint16_t), n, f);
    fwrite(B.data(), sizeof(uint16_t), n, f);
    fclose(f);
    printf("sr_probe: %d conversions written to %s\n", n, argv[1]);
    return 0;
}
