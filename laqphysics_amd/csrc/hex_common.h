// Shared device helpers for the laqphysics_hip engine (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

// complex128 as two doubles; matches numpy/torch complex128 memory.
struct __align__(16) zd {
    double x, y;
};

__host__ __device__ __forceinline__ zd zmake(double a, double b) { zd z; z.x = a; z.y = b; return z; }
__host__ __device__ __forceinline__ zd zadd(zd a, zd b) { return zmake(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ zd zsub(zd a, zd b) { return zmake(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ zd zmul(zd a, zd b) {
    return zmake(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ zd zscale(zd a, double s) { return zmake(a.x * s, a.y * s); }
__host__ __device__ __forceinline__ zd zconj(zd a) { return zmake(a.x, -a.y); }
// a * b + c as two fused multiply-adds per component (four FP64 operations;
// the form c + (a.x b.x - a.y b.y) would cost six)
__host__ __device__ __forceinline__ zd zfma(zd a, zd b, zd c) {
    return zmake(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
// conj(a) * b + c, likewise
__host__ __device__ __forceinline__ zd zcfma(zd a, zd b, zd c) {
    return zmake(fma(a.x, b.x, fma(a.y, b.y, c.x)), fma(a.x, b.y, fma(-a.y, b.x, c.y)));
}
// conj(a) * b
__host__ __device__ __forceinline__ zd zcmul(zd a, zd b) {
    return zmake(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
// a / b, Smith's algorithm (scaled to avoid overflow)
__host__ __device__ __forceinline__ zd zdiv(zd a, zd b) {
    if (fabs(b.x) >= fabs(b.y)) {
        double t = b.y / b.x, d = b.x + b.y * t;
        return zmake((a.x + a.y * t) / d, (a.y - a.x * t) / d);
    }
    double t = b.x / b.y, d = b.y + b.x * t;
    return zmake((a.x * t + a.y) / d, (a.y * t - a.x) / d);
}
__host__ __device__ __forceinline__ double zabs2(zd a) { return a.x * a.x + a.y * a.y; }

// Streaming store of one complex value (one 16-byte nontemporal store): for
// outputs that no later access of the same kernel reads, so they do not push
// the kernel's re-read inputs out of L2.
typedef double hex_v2d __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_stream(zd *p, zd v) {
    hex_v2d w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<hex_v2d *>(p));
}

// Deterministic wave64 reduction of a double (butterfly over DPP/shuffles).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

namespace hex {
void set_error(const std::string &msg);
}

#define HEX_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t hex_e_ = (call);                                                    \
        if (hex_e_ != hipSuccess) {                                                    \
            hex::set_error(std::string(#call) + ": " + hipGetErrorString(hex_e_));    \
            return HEX_EDEVICE;                                                        \
        }                                                                              \
    } while (0)
