// Shared device/host helpers for the gfx950 kernels: FP64 complex arithmetic on double2
// (interleaved re/im, the layout of ComplexF64), launch checks and small utilities.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

typedef double2 zd;   // complex FP64, 16-byte aligned -> global_load_dwordx4

__host__ __device__ __forceinline__ zd zmk(double re, double im) { zd r; r.x = re; r.y = im; return r; }
__host__ __device__ __forceinline__ zd zzero() { return zmk(0.0, 0.0); }
__host__ __device__ __forceinline__ zd zadd(zd a, zd b) { return zmk(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ zd zsub(zd a, zd b) { return zmk(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ zd zneg(zd a) { return zmk(-a.x, -a.y); }
__host__ __device__ __forceinline__ zd zconj(zd a) { return zmk(a.x, -a.y); }
__host__ __device__ __forceinline__ zd zmul(zd a, zd b) {
    return zmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__host__ __device__ __forceinline__ zd zcmul(zd a, zd b) {
    return zmk(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__host__ __device__ __forceinline__ zd zscale(zd a, double s) { return zmk(a.x * s, a.y * s); }
// c + a*b
__host__ __device__ __forceinline__ zd zfma(zd a, zd b, zd c) {
    return zmk(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
// c + conj(a)*b
__host__ __device__ __forceinline__ zd zcfma(zd a, zd b, zd c) {
    return zmk(fma(a.x, b.x, fma(a.y, b.y, c.x)), fma(a.x, b.y, fma(-a.y, b.x, c.y)));
}
__host__ __device__ __forceinline__ double zabs2(zd a) { return a.x * a.x + a.y * a.y; }
__host__ __device__ __forceinline__ double zabs(zd a) { return hypot(a.x, a.y); }
__host__ __device__ __forceinline__ zd zdiv(zd a, zd b) {
    // Smith's algorithm
    if (fabs(b.x) >= fabs(b.y)) {
        double r = b.y / b.x, d = b.x + r * b.y;
        return zmk((a.x + r * a.y) / d, (a.y - r * a.x) / d);
    } else {
        double r = b.x / b.y, d = b.y + r * b.x;
        return zmk((r * a.x + a.y) / d, (r * a.y - a.x) / d);
    }
}

#define RLS_WAVE 64

// Stores other workgroups read after a flag hand-off: relaxed atomic stores at agent scope
// are written through this XCD's L2 (sc1), so the release that publishes them has no dirty
// lines to write back (a plain store + agent-scope fence writes back the whole L2).
__device__ __forceinline__ void st_wt(double* p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_wt(zd* p, zd v) {
    double* q = reinterpret_cast<double*>(p);
    st_wt(q, v.x);
    st_wt(q + 1, v.y);
}

// Block-wide sum of one double per thread (blockDim.x multiple of 64, <= 1024).
__device__ __forceinline__ double block_sum(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    const int nw = (blockDim.x + 63) >> 6;
    double s = 0.0;
    for (int i = 0; i < nw; ++i) s += red[i];
    __syncthreads();
    return s;
}

// Two block-wide sums with one pair of barriers; red holds >= 2 * (blockDim.x / 64) doubles.
__device__ __forceinline__ void block_sum2(double& a, double& b, double* red) {
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) { red[2 * w] = a; red[2 * w + 1] = b; }
    __syncthreads();
    const int nw = (blockDim.x + 63) >> 6;
    double sa = 0.0, sb = 0.0;
    for (int i = 0; i < nw; ++i) { sa += red[2 * i]; sb += red[2 * i + 1]; }
    __syncthreads();
    a = sa;
    b = sb;
}

// block_sum / block_sum2 without the barriers that protect `red` against an earlier or a
// later use: the caller gives each reduction its own buffer and separates the uses of one
// buffer by other barriers.  Same sums in the same order.
__device__ __forceinline__ double block_sum_1b(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    const int nw = (blockDim.x + 63) >> 6;
    double s = 0.0;
    for (int i = 0; i < nw; ++i) s += red[i];
    return s;
}
__device__ __forceinline__ void block_sum2_1b(double& a, double& b, double* red) {
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { red[2 * w] = a; red[2 * w + 1] = b; }
    __syncthreads();
    const int nw = (blockDim.x + 63) >> 6;
    double sa = 0.0, sb = 0.0;
    for (int i = 0; i < nw; ++i) { sa += red[2 * i]; sb += red[2 * i + 1]; }
    a = sa;
    b = sb;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

struct RlsError {
    int code;
    std::string msg;
};

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            throw RlsError{-2, std::string(#expr) + ": " + hipGetErrorString(_e)};         \
    } while (0)

#define LAUNCH_CHECK()                                                                     \
    do {                                                                                   \
        hipError_t _e = hipGetLastError();                                                 \
        if (_e != hipSuccess)                                                              \
            throw RlsError{-2, std::string("kernel launch: ") + hipGetErrorString(_e)};    \
    } while (0)

static inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }
