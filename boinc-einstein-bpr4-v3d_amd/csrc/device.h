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

// The orbital offset of sample i in samples (src/science.cpp:36-40), rounding as the host
// rule brp_sample_offset does: the nested-FMA form with flag 2, else separate roundings.
__device__ __forceinline__ float sample_offset(uint32_t i, const brp_params& p, const float2* __restrict__ lut) {
    const float t = float(i) * p.dt;
    if (p.flags & 2u)
        return __builtin_fmaf(p.tau * fma_sin(__builtin_fmaf(p.omega, t, p.psi), lut), p.step_inv, -p.s0);
    return p.tau * lut_sin(p.omega * t + p.psi, lut) * p.step_inv - p.s0;
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
// RPI: the last step as one v_cvt_rpi_i32_f32 (floor(x + 0.5)) instead of the FP64 add and
// truncation.  The two agree for every float x in [-0.5, 2^24] (checked exhaustively on the
// device, 2.3e9 values); below -0.5 floor gives a negative index where the truncation can give
// 0, so a caller must treat every negative result as "use the exact form", and results of
// 2^24 or more likewise.
template <bool NONNEG, bool FMA, bool RPI = false>
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
    if constexpr (RPI) {
        int j;
        asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(j) : "v"(idx));
        return j;
    } else {
        return int(double(idx) + 0.5);
    }
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

// ------------------------------------------- split-FP16 MFMA DFT-16 (variant) --
// The precision variant runs a 16-point DFT stage as a real 32 x 32 GEMM on the matrix
// cores: rows i = 2k' + c' (output bin k', re/im), columns k = 2q + c (input q, re/im),
// F[2k'][2q] = F[2k'+1][2q+1] = cos t, F[2k'][2q+1] = sin t, F[2k'+1][2q] = -sin t with
// t = 2 pi (k' q mod 16) / 16.  Each wave holds 64 butterflies (one per lane); one
// v_mfma_f32_32x32x8_f16 takes 32 of them and 8 of the 32 rows of k.  Inputs and the
// matrix are split x = hi + lo into FP16 (hi = f16(x), lo = f16(x - hi)) and the products
// F_lo x_hi + F_hi x_lo + F_hi x_hi accumulate in FP32 (the F_lo x_lo term, ~2^-22
// relative, is dropped).  The wave's inputs are scaled by a power of two so that the
// largest magnitude is just below 2^14 (hi stays normal and finite; exact rescaling).
// Lane layout of the 32x32x8 instruction (checked on the device): A and B operands hold
// k = 4 (lane / 32) + e of row / column lane % 32, and accumulator register r holds row
// 8 (r / 4) + 4 (lane / 32) + r % 4 of column lane % 32.  v_permlane32_swap moves the
// halves between lane j and lane j + 32.
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

struct Dft16Mfma {
    h4 ahi[4], alo[4];  // this lane's A operand of each k block, hi and lo parts
};

__device__ __forceinline__ float dft16_entry(int i, int k) {
    constexpr float kCos[16] = {1.0f, kC16, kS8, kS16, 0.0f, -kS16, -kS8, -kC16,
                                -1.0f, -kC16, -kS8, -kS16, 0.0f, kS16, kS8, kC16};
    const int t = ((i >> 1) * (k >> 1)) & 15;
    const float c = kCos[t], s = kCos[(t + 12) & 15];  // sin(x) = cos(x - pi/2)
    const int cc = ((i & 1) << 1) | (k & 1);
    return cc == 0 || cc == 3 ? c : (cc == 1 ? s : -s);
}

__device__ __forceinline__ Dft16Mfma dft16_mfma_setup() {
    const int lane = int(threadIdx.x & 63u);
    Dft16Mfma f;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float v = dft16_entry(lane & 31, 8 * kk + 4 * (lane >> 5) + e);
            const _Float16 hi = _Float16(v);
            f.ahi[kk][e] = hi;
            f.alo[kk][e] = _Float16(v - float(hi));
        }
    return f;
}

__device__ __forceinline__ void swap_halves(h4& a, h4& b) {
    uint32_t ua[2], ub[2];
    __builtin_memcpy(ua, &a, 8);
    __builtin_memcpy(ub, &b, 8);
#pragma unroll
    for (int w = 0; w < 2; ++w) {
        const auto r = __builtin_amdgcn_permlane32_swap(ua[w], ub[w], false, false);
        ua[w] = r[0];
        ub[w] = r[1];
    }
    __builtin_memcpy(&a, ua, 8);
    __builtin_memcpy(&b, ub, 8);
}

// Wave-collective: every lane of the wave calls it with its butterfly in v[0..15];
// returns the DFT in natural order.
__device__ __forceinline__ void dft16_mfma(cf* v, const Dft16Mfma& f) {
    float mx = 0.0f;
#pragma unroll
    for (int q = 0; q < 16; ++q) mx = fmaxf(mx, fmaxf(fabsf(v[q].x), fabsf(v[q].y)));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    int ex = 0;
    (void)frexpf(mx, &ex);                          // mx < 2^ex
    const float scale = mx > 0.0f ? ldexpf(1.0f, 14 - ex) : 1.0f;
    const float unscale = mx > 0.0f ? ldexpf(1.0f, ex - 14) : 1.0f;
    f32x16 da = {}, db = {};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        h4 phi, plo, qhi, qlo;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const cf xp = v[4 * kk + (e >> 1)], xq = v[4 * kk + 2 + (e >> 1)];
            const float a = ((e & 1) ? xp.y : xp.x) * scale, b = ((e & 1) ? xq.y : xq.x) * scale;
            phi[e] = _Float16(a);
            plo[e] = _Float16(a - float(phi[e]));
            qhi[e] = _Float16(b);
            qlo[e] = _Float16(b - float(qhi[e]));
        }
        swap_halves(phi, qhi);  // lanes < 32 now feed butterflies 0..31 through p, 32..63 through q
        swap_halves(plo, qlo);
        da = __builtin_amdgcn_mfma_f32_32x32x8f16(f.alo[kk], phi, da, 0, 0, 0);
        da = __builtin_amdgcn_mfma_f32_32x32x8f16(f.ahi[kk], plo, da, 0, 0, 0);
        da = __builtin_amdgcn_mfma_f32_32x32x8f16(f.ahi[kk], phi, da, 0, 0, 0);
        db = __builtin_amdgcn_mfma_f32_32x32x8f16(f.alo[kk], qhi, db, 0, 0, 0);
        db = __builtin_amdgcn_mfma_f32_32x32x8f16(f.ahi[kk], qlo, db, 0, 0, 0);
        db = __builtin_amdgcn_mfma_f32_32x32x8f16(f.ahi[kk], qhi, db, 0, 0, 0);
    }
    // After the swap every lane holds its own butterfly: rows with even (i / 4) in da,
    // odd in db, register 4 (i / 8) + i % 4.
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const auto s = __builtin_amdgcn_permlane32_swap(__float_as_uint(da[r]), __float_as_uint(db[r]), false, false);
        da[r] = __uint_as_float(s[0]);
        db[r] = __uint_as_float(s[1]);
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int i = 2 * k;  // re row i, im row i + 1 (same quad)
        const int r = 4 * (i >> 3) + (i & 3);
        const bool odd = (i >> 2) & 1;
        v[k] = cf{(odd ? db[r] : da[r]) * unscale, (odd ? db[r + 1] : da[r + 1]) * unscale};
    }
}

}  // namespace brp_dev
