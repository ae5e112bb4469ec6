// common.hpp - device helpers shared by every kernel of libmctdhb_hip.
//
// Complex numbers are interleaved (re, im) pairs, double2 for FP64 and
// float2 for the mixed-precision forms.  All kernels are written for gfx950
// (wave64, 160 KB LDS per CU, 8 XCDs).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

namespace mctdhb {

// LDS one workgroup may allocate on gfx950 (MI355X): 160 KiB.
constexpr size_t kLdsPerWorkgroup = 160 * 1024;

// ------------------------------------------------------------ error model
// Every C-ABI entry point returns an int status (0 = ok) and leaves a
// message in a thread-local buffer (mctdhb_last_error()).
enum Status : int {
    kOk = 0,
    kInvalidArgument = 1,
    kOverflow = 2,
    kDeviceError = 3,
    kOutOfMemory = 4,
    kNotBuilt = 5,
    kNumerical = 6,
};

void set_error(const std::string& msg);
const char* last_error();

#define MCTDHB_HIP_CHECK(expr)                                                            \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess) {                                                           \
            ::mctdhb::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));       \
            return ::mctdhb::kDeviceError;                                                \
        }                                                                                 \
    } while (0)

// ------------------------------------------------------------ complex math
template <typename R> struct cplx_t;
template <> struct cplx_t<double> { using type = double2; };
template <> struct cplx_t<float> { using type = float2; };
template <typename R> using cplx = typename cplx_t<R>::type;

template <typename R> __host__ __device__ __forceinline__ cplx<R> mk(R re, R im) {
    cplx<R> z;
    z.x = re;
    z.y = im;
    return z;
}
__host__ __device__ __forceinline__ double2 operator+(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ double2 operator-(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ float2 operator+(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ float2 operator-(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

template <typename Z> __host__ __device__ __forceinline__ Z cmul(Z a, Z b) {
    Z r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
// a * conj(b)... conj(a) * b, the inner-product kernel
template <typename Z> __host__ __device__ __forceinline__ Z cconjmul(Z a, Z b) {
    Z r;
    r.x = a.x * b.x + a.y * b.y;
    r.y = a.x * b.y - a.y * b.x;
    return r;
}
// acc + a * b (complex)
template <typename Z> __host__ __device__ __forceinline__ Z cfma(Z a, Z b, Z acc) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
    return acc;
}
template <typename Z, typename R> __host__ __device__ __forceinline__ Z cscale(Z a, R s) {
    a.x *= s;
    a.y *= s;
    return a;
}

__device__ __forceinline__ double2 to_d2(double2 a) { return a; }
__device__ __forceinline__ double2 to_d2(float2 a) { return make_double2((double) a.x, (double) a.y); }
template <typename Z> __device__ __forceinline__ Z from_d2(double2 a);
template <> __device__ __forceinline__ double2 from_d2<double2>(double2 a) { return a; }
template <> __device__ __forceinline__ float2 from_d2<float2>(double2 a) { return make_float2((float) a.x, (float) a.y); }

// ----------------------------------------------------------- wave reduce
// lane i gets v[i - K] within its row of 16 (0 where that leaves the row);
// every lane of the wave must be active
template <int K> __device__ __forceinline__ double row_shr(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x110 + K, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x110 + K, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
// Lane 15 of each row of 16 gets the row's sum (fixed order); every lane of
// the wave must be active.
__device__ __forceinline__ double row_sum(double v) {
    v += row_shr<1>(v);
    v += row_shr<2>(v);
    v += row_shr<4>(v);
    return v + row_shr<8>(v);
}

__device__ __forceinline__ double readlane_d(double v, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}
// Sum of v over the 64 lanes of the wave, the same value in every lane, in a
// fixed order: DPP row sums (vector ALU), then the four row totals through
// readlane (cross-lane shuffles are LDS permutes whose dependent chain cost
// each sum some 1000 cycles).  Every lane of the wave must be active.
__device__ __forceinline__ double wave_sum(double v) {
    v = row_sum(v);
    return (readlane_d(v, 15) + readlane_d(v, 31)) + (readlane_d(v, 47) + readlane_d(v, 63));
}


// Block-wide sum of NV doubles per thread into lane-0-of-block results.
// `scratch` must hold NV * (blockDim.x / 64) doubles.  Returns the totals in
// thread 0 (other threads get garbage).  Deterministic for a fixed launch.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* scratch) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int t = 0; t < NV; t++) v[t] = wave_sum(v[t]);
    if (lane == 0) {
#pragma unroll
        for (int t = 0; t < NV; t++) scratch[t * nw + wid] = v[t];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int t = 0; t < NV; t++) {
            double s = 0;
            for (int w = 0; w < nw; w++) s += scratch[t * nw + w];
            v[t] = s;
        }
    }
    __syncthreads();
}

// ------------------------------------------------------ XCD-aware blocks
// The hardware deals consecutive workgroups round-robin over the 8 XCDs.
// Remapping lets each XCD own a contiguous range of configurations, so the
// coefficient windows its gathers touch stay in that XCD's L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t bid, uint32_t nblocks) {
    constexpr uint32_t kXcd = 8;
    const uint32_t per = nblocks / kXcd, rem = nblocks % kXcd;
    const uint32_t x = bid % kXcd, r = bid / kXcd;
    // XCD x owns per (+1 for the first rem XCDs) consecutive blocks
    const uint32_t start = x * per + (x < rem ? x : rem);
    return start + r;
}

inline unsigned div_up(uint64_t a, uint64_t b) { return (unsigned) ((a + b - 1) / b); }

}  // namespace mctdhb
