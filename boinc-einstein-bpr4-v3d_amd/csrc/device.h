// device.h -- gfx950 device helpers shared by the plain and fused engines:
// the bit-exact resampling index, complex arithmetic, small DFT kernels and the
// dyadic harmonic sums.  Every file including this one is compiled with
// -ffp-contract=off; the FP32 resampling expressions below must round exactly as
// written to reproduce the host rule (src/science.cpp:20-40, 75-85).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "brp_hip.h"

#pragma clang fp contract(off)

namespace brp_dev {

struct cf { float x, y; };

__device__ __forceinline__ cf make(float x, float y) { return cf{x, y}; }
__device__ __forceinline__ cf add(cf a, cf b) { return cf{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cf sub(cf a, cf b) { return cf{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cf scale(cf a, float s) { return cf{a.x * s, a.y * s}; }
__device__ __forceinline__ cf conj(cf a) { return cf{a.x, -a.y}; }
// a * b with one rounding per component after the product pair (FMA form).
__device__ __forceinline__ cf mul(cf a, cf b) {
    return cf{__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x)};
}
// -i * a
__device__ __forceinline__ cf mul_negi(cf a) { return cf{a.y, -a.x}; }
__device__ __forceinline__ cf load(const float2* p) { float2 v = *p; return cf{v.x, v.y}; }
__device__ __forceinline__ void store(float2* p, cf v) { *p = make_float2(v.x, v.y); }

// ------------------------------------------------------------- resampling --
// FP32 LUT sine, separately rounded (src/science.cpp:20-28).  lut holds interleaved
// (sin, cos) pairs for 65 segments.
__device__ __forceinline__ float lut_sin(float phase, const float2* __restrict__ lut) {
    const float scaled = (1.0f / 6.283185f) * phase;
    float frac = scaled - __builtin_truncf(scaled);
    if (frac < 0.0f) frac += 1.0f;
    const int seg = int(frac * 64.0f + 0.5f);
    const float d = 6.283185f * (frac - 0.015625f * float(seg));
    const float d2 = d * (0.5f * d);
    const float2 sc = lut[seg];
    return sc.x + d * sc.y - d2 * sc.x;
}

// Nested-FMA LUT sine for flag 2 (src/science.cpp:29-35).
__device__ __forceinline__ float fma_sin(float phase, const float2* __restrict__ lut) {
    const float scaled = (1.0f / 6.283185f) * phase;
    float frac = scaled - __builtin_truncf(scaled);
    if (frac < 0.0f) frac += 1.0f;
    const int seg = int(__builtin_fmaf(frac, 64.0f, 0.5f));
    const float d = 6.283185f * __builtin_fmaf(-float(seg), 0.015625f, frac);
    const float d2 = (0.5f * d) * d;
    const float2 sc = lut[seg];
    return __builtin_fmaf(-d2, sc.x, __builtin_fmaf(d, sc.y, sc.x));
}

struct Orbit32 { float tau, omega, psi, dt, step_inv, s0; uint32_t flags; };

// Source index of resampled sample i: int(double(float(i) - offset(i)) + 0.5), the
// CPU rule of src/science.cpp:80 (truncation toward zero after the double add).
__device__ __forceinline__ int source_index(uint32_t i, const Orbit32& o, const float2* __restrict__ lut) {
    const float t = float(i) * o.dt;
    float off;
    if (o.flags & 2u)
        off = __builtin_fmaf(o.tau * fma_sin(__builtin_fmaf(o.omega, t, o.psi), lut), o.step_inv, -o.s0);
    else
        off = o.tau * lut_sin(o.omega * t + o.psi, lut) * o.step_inv - o.s0;
    const float idx = float(i) - off;
    return int(double(idx) + 0.5);
}

// The same index with fewer instructions, rounding identically:
//  * frac * 64 and seg / 64 are exact, so fl(fl(frac * 64) + 0.5) = fma(frac, 64, 0.5) and
//    fl(frac - fl(seg / 64)) = fma(-seg, 1/64, frac) in both sine forms;
//  * d / 2 is exact, so d2 = fl(d * fl(d / 2)) = fl(d * d) / 2 and fl(d2 * sin) =
//    fl(fl(d * d) * (sin / 2)); lut4 holds (sin, cos, sin / 2, 0) per segment;
//  * NONNEG (phase >= 0 for every sample: omega, psi, dt >= 0): trunc = floor, the
//    wrap-around never fires, and frac is the exact hardware fract.
// Checked bit for bit against lut_sin / fma_sin on the CPU for 8e8 phases, including
// subnormals and phases at segment centres.
// FMA: flag 2 (the caller branches on it once per block).
template <bool NONNEG, bool FMA>
__device__ __forceinline__ int source_index4(uint32_t i, const Orbit32& o, const float4* __restrict__ lut4) {
    const float fi = float(i);
    const float t = fi * o.dt;
    const float phase = FMA ? __builtin_fmaf(o.omega, t, o.psi) : o.omega * t + o.psi;
    const float scaled = (1.0f / 6.283185f) * phase;
    float frac;
    if constexpr (NONNEG) {
        frac = __builtin_amdgcn_fractf(scaled);
    } else {
        frac = scaled - __builtin_truncf(scaled);
        if (frac < 0.0f) frac += 1.0f;
    }
    const int seg = int(__builtin_fmaf(frac, 64.0f, 0.5f));
    const float d = 6.283185f * __builtin_fmaf(-float(seg), 0.015625f, frac);
    const float dd = d * d;
    const float4 sc = lut4[seg];
    float off;
    if constexpr (FMA) {
        const float s = __builtin_fmaf(-dd, sc.z, __builtin_fmaf(d, sc.y, sc.x));
        off = __builtin_fmaf(o.tau * s, o.step_inv, -o.s0);
    } else {
        const float s = (sc.x + d * sc.y) - dd * sc.z;
        off = o.tau * s * o.step_inv - o.s0;
    }
    const float idx = fi - off;
    return int(double(idx) + 0.5);
}

// sin(x) in FP64 for orbital phases (|x| well below 2^30): Cody-Waite reduction by
// multiples of pi/2 (pi/2 split into a double head and tail) to |r| <= pi/4, then the
// Taylor series of sin / cos through r^17 / r^16 in Horner form; the first omitted term
// is below 1e-17 on that interval.
__device__ __forceinline__ double sin_f64(double x) {
    const double k = __builtin_rint(x * 0.63661977236758134308);
    const double r = __builtin_fma(-k, 6.123233995736766036e-17, __builtin_fma(-k, 1.5707963267948966192, x));
    const double r2 = r * r;
    double s = 2.8114572543455207632e-15;                 //  1/17!
    s = __builtin_fma(s, r2, -7.6471637318198164759e-13);  // -1/15!
    s = __builtin_fma(s, r2, 1.6059043836821614599e-10);   //  1/13!
    s = __builtin_fma(s, r2, -2.5052108385441718775e-8);   // -1/11!
    s = __builtin_fma(s, r2, 2.7557319223985890653e-6);    //  1/9!
    s = __builtin_fma(s, r2, -1.9841269841269841270e-4);   // -1/7!
    s = __builtin_fma(s, r2, 8.3333333333333333333e-3);    //  1/5!
    s = __builtin_fma(s, r2, -1.6666666666666666667e-1);   // -1/3!
    s = __builtin_fma(s * r2, r, r);
    double c = 4.7794773323873852974e-14;                 //  1/16!
    c = __builtin_fma(c, r2, -1.1470745597729724714e-11);  // -1/14!
    c = __builtin_fma(c, r2, 2.0876756987868098979e-9);    //  1/12!
    c = __builtin_fma(c, r2, -2.7557319223985890653e-7);   // -1/10!
    c = __builtin_fma(c, r2, 2.4801587301587301587e-5);    //  1/8!
    c = __builtin_fma(c, r2, -1.3888888888888888889e-3);   // -1/6!
    c = __builtin_fma(c, r2, 4.1666666666666666667e-2);    //  1/4!
    c = __builtin_fma(c, r2, -0.5);                        // -1/2!
    c = __builtin_fma(c, r2, 1.0);
    const int q = int(int64_t(k) & 3);
    return q == 0 ? s : (q == 1 ? c : (q == 2 ? -s : -c));
}

// FP64 variant: the orbital offset from the FP64 sine, rounded to the nearest sample.
__device__ __forceinline__ int source_index_fp64(uint32_t i, double tau_dt, double omega, double psi,
                                                 double dt, double s0_dt) {
    const double t = double(i) * dt;
    const double off = tau_dt * sin_f64(omega * t + psi) - s0_dt;
    return int(floor(double(i) - off + 0.5));
}

// ----------------------------------------------------------------- DFTs --
// Forward DFTs of size 2, 3, 4, 8, 16 on registers, natural-order output.
constexpr float kS8 = 0.70710678118654752440f;   // sqrt(1/2)
constexpr float kC16 = 0.92387953251128675613f;  // cos(pi/8)
constexpr float kS16 = 0.38268343236508977173f;  // sin(pi/8)
constexpr float kS3 = 0.86602540378443864676f;   // sin(2pi/3)

__device__ __forceinline__ void dft2(cf& a, cf& b) { cf t = a; a = add(t, b); b = sub(t, b); }

__device__ __forceinline__ void dft3(cf& a, cf& b, cf& c) {
    const cf s = add(b, c), d = sub(b, c);
    const cf m = sub(a, scale(s, 0.5f));
    const cf r = cf{d.y * kS3, -d.x * kS3};  // -i sin(2pi/3) d
    a = add(a, s);
    b = add(m, r);
    c = sub(m, r);
}

__device__ __forceinline__ void dft4(cf& a, cf& b, cf& c, cf& d) {
    const cf s02 = add(a, c), d02 = sub(a, c), s13 = add(b, d), d13 = mul_negi(sub(b, d));
    a = add(s02, s13);
    b = add(d02, d13);
    c = sub(s02, s13);
    d = sub(d02, d13);
}

template <int R>
__device__ __forceinline__ void dft(cf* v);

template <> __device__ __forceinline__ void dft<2>(cf* v) { dft2(v[0], v[1]); }
template <> __device__ __forceinline__ void dft<3>(cf* v) { dft3(v[0], v[1], v[2]); }
template <> __device__ __forceinline__ void dft<4>(cf* v) { dft4(v[0], v[1], v[2], v[3]); }

template <> __device__ __forceinline__ void dft<8>(cf* v) {
    cf e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    cf o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3);
    dft4(o0, o1, o2, o3);
    o1 = cf{(o1.x + o1.y) * kS8, (o1.y - o1.x) * kS8};     // * e^{-i pi/4}
    o2 = mul_negi(o2);                                      // * -i
    o3 = cf{(o3.y - o3.x) * kS8, -(o3.x + o3.y) * kS8};    // * e^{-3i pi/4}
    v[0] = add(e0, o0); v[4] = sub(e0, o0);
    v[1] = add(e1, o1); v[5] = sub(e1, o1);
    v[2] = add(e2, o2); v[6] = sub(e2, o2);
    v[3] = add(e3, o3); v[7] = sub(e3, o3);
}

// 16 = 4 x 4: DFT4 over n2 of x[n1 + 4 n2], twiddle W16^{n1 k2}, DFT4 over n1;
// output index k2 + 4 k1.
template <> __device__ __forceinline__ void dft<16>(cf* v) {
    cf a[4][4];
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) {
        a[n1][0] = v[n1]; a[n1][1] = v[n1 + 4]; a[n1][2] = v[n1 + 8]; a[n1][3] = v[n1 + 12];
        dft4(a[n1][0], a[n1][1], a[n1][2], a[n1][3]);
    }
    // W16^1, W16^2, W16^3, W16^4 = -i, W16^6, W16^9
    const cf w1 = cf{kC16, -kS16}, w2 = cf{kS8, -kS8}, w3 = cf{kS16, -kC16};
    const cf w6 = cf{-kS8, -kS8}, w9 = cf{-kC16, kS16};
    a[1][1] = mul(a[1][1], w1); a[1][2] = mul(a[1][2], w2); a[1][3] = mul(a[1][3], w3);
    a[2][1] = mul(a[2][1], w2); a[2][2] = mul_negi(a[2][2]); a[2][3] = mul(a[2][3], w6);
    a[3][1] = mul(a[3][1], w3); a[3][2] = mul(a[3][2], w6); a[3][3] = mul(a[3][3], w9);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
        cf b0 = a[0][k2], b1 = a[1][k2], b2 = a[2][k2], b3 = a[3][k2];
        dft4(b0, b1, b2, b3);
        v[k2] = b0; v[k2 + 4] = b1; v[k2 + 8] = b2; v[k2 + 12] = b3;
    }
}

// ------------------------------------------------------- harmonic sums --
// Dyadic FP32 sums of one bin in the historical order (shaders/hs_sum.glsl:2-10):
// s[h] is the 2^h-harmonic sum, index (i*k + 8) >> 4 = round-half-up(i*k/16).
__device__ __forceinline__ void dyadic_sums(const float* __restrict__ P, uint32_t i, float s[5]) {
#define BRP_AT(k) P[(i * (k) + 8u) >> 4]
    float v = P[i];
    s[0] = v;
    v += BRP_AT(8); s[1] = v;
    v += BRP_AT(12) + BRP_AT(4); s[2] = v;
    v += BRP_AT(14) + BRP_AT(10) + BRP_AT(6) + BRP_AT(2); s[3] = v;
    v += BRP_AT(15) + BRP_AT(13) + BRP_AT(11) + BRP_AT(9) + BRP_AT(7) + BRP_AT(5) + BRP_AT(3) + BRP_AT(1);
    s[4] = v;
#undef BRP_AT
}

// Result slot: 4-word header {count, flags, 0, 0} then capacity candidates of 4 words
// (the layout of shaders/common.glsl:23-30).  Flag 1 = candidate overflow, 2 = a
// resample index outside the observation.
__device__ __forceinline__ void append_candidate(uint32_t* slot, uint32_t capacity, uint32_t bin,
                                                 uint32_t level, float power) {
    const uint32_t at = atomicAdd(slot, 1u);
    if (at < capacity) {
        uint4* c = reinterpret_cast<uint4*>(slot + 4) + at;
        *c = make_uint4(bin, level, __float_as_uint(power), 0u);
    } else {
        atomicOr(slot + 1, 1u);
    }
}

}  // namespace brp_dev
