// Runs gfx950's stochastic-rounding conversions (v_cvt_sr_f16_f32,
// v_cvt_sr_bf16_f32) over a table of (binary32 bits, random word) pairs
// read from a text file and writes the results next to them.  Used by
// make_sr_table.py to record the hardware rule the oracle restates.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <vector>

typedef __fp16 half2_t __attribute__((ext_vector_type(2)));
typedef short short2_t __attribute__((ext_vector_type(2)));

__global__ void k_sr(const float *x, const uint32_t *r, uint32_t *oh, uint32_t *ob, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const half2_t h = __builtin_amdgcn_cvt_sr_f16_f32(half2_t{}, x[t], r[t], false);
    oh[t] = __builtin_bit_cast(uint32_t, h) & 0xffffu;
    const short2_t b = __builtin_amdgcn_cvt_sr_bf16_f32(short2_t{}, x[t], r[t], false);
    ob[t] = __builtin_bit_cast(uint32_t, b) & 0xffffu;
}

#define CHK(e)                                                                   \
    do {                                                                         \
        hipError_t err_ = (e);                                                   \
        if (err_ != hipSuccess) {                                                \
            fprintf(stderr, "%s: %s\n", #e, hipGetErrorString(err_));            \
            return 1;                                                            \
        }                                                                        \
    } while (0)

int main(int argc, char **argv)
{
    if (argc != 3) return 2;
    FILE *f = fopen(argv[1], "r");
    if (!f) return 2;
    std::vector<float> x;
    std::vector<uint32_t> r;
    unsigned a, b;
    while (fscanf(f, "%x %x", &a, &b) == 2) {
        float v;
        memcpy(&v, &a, 4);
        x.push_back(v);
        r.push_back(b);
    }
    fclose(f);
    const int n = (int)x.size();
    if (n == 0) return 2;
    float *dx;
    uint32_t *dr, *dh, *db;
    CHK(hipMalloc(&dx, 4 * (size_t)n));
    CHK(hipMalloc(&dr, 4 * (size_t)n));
    CHK(hipMalloc(&dh, 4 * (size_t)n));
    CHK(hipMalloc(&db, 4 * (size_t)n));
    CHK(hipMemcpy(dx, x.data(), 4 * (size_t)n, hipMemcpyHostToDevice));
    CHK(hipMemcpy(dr, r.data(), 4 * (size_t)n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_sr, dim3((n + 255) / 256), dim3(256), 0, 0, dx, dr, dh, db, n);
    CHK(hipGetLastError());
    std::vector<uint32_t> h(n), bb(n);
    CHK(hipMemcpy(h.data(), dh, 4 * (size_t)n, hipMemcpyDeviceToHost));
    CHK(hipMemcpy(bb.data(), db, 4 * (size_t)n, hipMemcpyDeviceToHost));
    FILE *o = fopen(argv[2], "w");
    if (!o) return 2;
    for (int i = 0; i < n; i++) fprintf(o, "%08x %08x %04x %04x\n", __builtin_bit_cast(uint32_t, x[i]), r[i], h[i], bb[i]);
    fclose(o);
    printf("sr_probe: %d conversions\n", n);
    return 0;
}
