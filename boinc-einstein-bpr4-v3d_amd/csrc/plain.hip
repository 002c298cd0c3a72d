// plain.hip -- the plain engine: one template per launch sequence, one kernel per stage,
// the dispatch structure of the reference's recorded command buffer
// (src/vulkan.cpp:221-266): resample -> mean -> pack -> Stockham passes -> radix-3 ->
// real split (+ power) -> power -> harmonic (gather, or init/scatter/collect).
//
// Each kernel is written for gfx950 wave64: 256-thread blocks, FP64 block sums for the
// pad mean, global Stockham radix-2/4/8 passes with twiddles from one table of
// W_fft_n^t computed in double.  The fused engine (fused.hip) is the fast path; this one
// is the baseline "one template per launch" GPU path and an independent cross-check.
#include <stdexcept>
#include <string>

#include "device.h"
#include "engine.h"

#pragma clang fp contract(off)

using namespace brp_dev;

namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void k_clear_slots(uint32_t* results, uint32_t slot_words, uint32_t nslots) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nslots) {
        uint32_t* h = results + size_t(s) * slot_words;
        h[0] = 0; h[1] = 0; h[2] = 0; h[3] = 0;
    }
}

// shaders/resample.comp:7-29: gather raw[i] = x[j(i)] for i < steps, flag an index
// outside the observation, and write one FP64 partial sum per block for the pad mean.
__global__ __launch_bounds__(kBlock) void k_resample(const float* __restrict__ x, const float2* __restrict__ lut,
                                                     brp_params p, float* __restrict__ raw,
                                                     int32_t* __restrict__ index, double* __restrict__ partial,
                                                     uint32_t* __restrict__ slot) {
    __shared__ double sums[kBlock / 64];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    double v = 0.0;
    if (i < p.steps) {
        const Orbit32 o{p.tau, p.omega, p.psi, p.dt, p.step_inv, p.s0, p.flags};
        const int j = source_index(i, o, lut);
        float y = 0.0f;
        if (j < 0 || uint32_t(j) >= p.unpadded) atomicOr(slot + 1, 2u);
        else y = x[j];
        raw[i] = y;
        if (index) index[i] = j;
        v = double(y);
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) sums[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) t += sums[w];
        partial[blockIdx.x] = t;
    }
}

// shaders/mean.comp:5-15, in FP64 with a fixed reduction order.
__global__ __launch_bounds__(kBlock) void k_mean(const double* __restrict__ partial, uint32_t blocks,
                                                 uint32_t steps, float* __restrict__ mean) {
    __shared__ double sums[kBlock];
    double v = 0.0;
    for (uint32_t b = threadIdx.x; b < blocks; b += kBlock) v += partial[b];
    sums[threadIdx.x] = v;
    __syncthreads();
    for (int d = kBlock / 2; d > 0; d >>= 1) {
        if (int(threadIdx.x) < d) sums[threadIdx.x] += sums[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) *mean = float(sums[0] / double(steps));
}

// shaders/prepare.comp:4-13 (Stockham order, no bit reversal): pack reals into complex
// pairs (packed) or take them as real parts, padding past `steps` with the mean.
__global__ __launch_bounds__(kBlock) void k_pack(const float* __restrict__ raw, const float* __restrict__ mean,
                                                 uint32_t steps, uint32_t fft_n, uint32_t packed,
                                                 float2* __restrict__ a) {
    const uint32_t m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= fft_n) return;
    const float mu = *mean;
    if (packed) {
        const uint32_t i = 2 * m;
        a[m] = make_float2(i < steps ? raw[i] : mu, i + 1 < steps ? raw[i + 1] : mu);
    } else {
        a[m] = make_float2(m < steps ? raw[m] : mu, 0.0f);
    }
}

// One global Stockham pass of radix R over a length-len transform:
// v_q = x[j + q len/R] * W_{ns R}^{(j mod ns) q}, y[(j/ns) ns R + (j mod ns) + q ns] = DFT_R(v)_q.
// tw[t] = W_len^t, so W_{ns R}^{kq} = tw[k q len/(ns R)].
template <int R>
__global__ __launch_bounds__(kBlock) void k_stockham(const float2* __restrict__ x, float2* __restrict__ y,
                                                     const float2* __restrict__ tw, uint32_t len, uint32_t ns) {
    const uint32_t span = len / R;
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= span) return;
    const uint32_t k = j % ns;
    const uint32_t step = len / (ns * R);
    cf v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = load(x + j + q * span);
    if (ns > 1) {
#pragma unroll
        for (int q = 1; q < R; ++q) v[q] = mul(v[q], load(tw + size_t(k) * q * step));
    }
    dft<R>(v);
    const uint32_t out = (j / ns) * ns * R + k;
#pragma unroll
    for (int q = 0; q < R; ++q) store(y + out + q * ns, v[q]);
}

// shaders/fft_real.comp:6-14: X[k] = 1/2 [(Z_k + conj Z_{M-k}) - i W_n^k (Z_k - conj Z_{M-k})],
// k = 0..M, with the power P[k] = |X_k|^2 / n fused when requested.
__global__ __launch_bounds__(kBlock) void k_real(const float2* __restrict__ z, const float2* __restrict__ twr,
                                                 uint32_t m, float inv_n, float2* __restrict__ X,
                                                 float* __restrict__ power) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k > m) return;
    const cf zk = load(z + (k == m ? 0 : k));
    const cf zr = conj(load(z + (k == 0 ? 0 : m - k)));
    const cf s = add(zk, zr), d = mul(sub(zk, zr), load(twr + k));
    const cf r = scale(add(s, mul_negi(d)), 0.5f);
    store(X + k, r);
    if (power && k < m) power[k] = k == 0 ? 0.0f : (r.x * r.x + r.y * r.y) * inv_n;
}

// shaders/power.comp:4-8
__global__ __launch_bounds__(kBlock) void k_power(const float2* __restrict__ X, uint32_t bins, float inv_n,
                                                  float* __restrict__ power) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= bins) return;
    const float2 v = X[k];
    power[k] = k == 0 ? 0.0f : (v.x * v.x + v.y * v.y) * inv_n;
}

// shaders/harmonic.comp:14-36: gather form, one thread per (fundamental j, level h).
__global__ __launch_bounds__(kBlock) void k_harmonic(const float* __restrict__ P, float* __restrict__ summed,
                                                     brp_params p, uint32_t* __restrict__ slot) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t h = blockIdx.y;
    if (j >= p.fundamental) return;
    float value = 0.0f;
    if (h == 0) {
        value = P[j];
    } else {
        const int64_t width = int64_t(1) << h;
        const int64_t begin = max(int64_t(p.window), int64_t(j) * width - width / 2);
        const int64_t end = min(int64_t(p.harmonic), int64_t(j) * width + width / 2);
        if (begin < end) {
            float s[5];
            dyadic_sums(P, uint32_t(begin), s);
            const float first = s[h];
            float best = first;
            for (int64_t i = begin + 1; i < end; ++i) {
                dyadic_sums(P, uint32_t(i), s);
                best = fmaxf(best, s[h]);
            }
            value = best > p.threshold[h] ? best : first;
        }
    }
    summed[size_t(h) * p.fundamental + j] = value;
    if (j >= p.window && value > p.threshold[h]) append_candidate(slot, p.capacity, j, h, value);
}

// shaders/hs_init.comp:6-18: every cell starts at its first in-window sum.
__global__ __launch_bounds__(kBlock) void k_hs_init(const float* __restrict__ P, float* __restrict__ summed,
                                                    brp_params p) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t h = blockIdx.y;
    if (j >= p.fundamental) return;
    float value = 0.0f;
    if (h == 0) {
        value = P[j];
    } else {
        const int64_t width = int64_t(1) << h;
        const int64_t begin = max(int64_t(p.window), int64_t(j) * width - width / 2);
        const int64_t end = min(int64_t(p.harmonic), int64_t(j) * width + width / 2);
        if (begin < end) {
            float s[5];
            dyadic_sums(P, uint32_t(begin), s);
            value = s[h];
        }
    }
    summed[size_t(h) * p.fundamental + j] = value;
}

// shaders/hs_scatter.comp:7-16: raise a cell to any above-threshold sum of its range.
// Sums are non-negative, so the unsigned order of the bits is the float order.
__global__ __launch_bounds__(kBlock) void k_hs_scatter(const float* __restrict__ P, float* __restrict__ summed,
                                                       brp_params p) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x + p.window;
    if (i >= p.harmonic) return;
    float s[5];
    dyadic_sums(P, i, s);
#pragma unroll
    for (uint32_t h = 1; h <= 4; ++h) {
        const uint32_t j = (i + (1u << (h - 1))) >> h;
        if (j < p.fundamental && s[h] > p.threshold[h])
            atomicMax(reinterpret_cast<uint32_t*>(summed) + size_t(h) * p.fundamental + j, __float_as_uint(s[h]));
    }
}

// shaders/hs_collect.comp:4-13
__global__ __launch_bounds__(kBlock) void k_hs_collect(const float* __restrict__ summed, brp_params p,
                                                       uint32_t* __restrict__ slot) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t h = blockIdx.y;
    if (j >= p.fundamental || j < p.window) return;
    const float value = summed[size_t(h) * p.fundamental + j];
    if (value > p.threshold[h]) append_candidate(slot, p.capacity, j, h, value);
}

uint32_t blocks(uint64_t n) { return uint32_t((n + kBlock - 1) / kBlock); }

}  // namespace

namespace brp_engine {

void plain_clear_slots(hipStream_t st, uint32_t* results, uint32_t slot_words, uint32_t nslots) {
    k_clear_slots<<<blocks(nslots), kBlock, 0, st>>>(results, slot_words, nslots);
}

void plain_resample(hipStream_t st, const PlainBuffers& b, const brp_params& p, uint32_t* slot) {
    k_resample<<<blocks(p.steps), kBlock, 0, st>>>(b.x, b.lut, p, b.raw, b.index, b.partial, slot);
    k_mean<<<1, kBlock, 0, st>>>(b.partial, blocks(p.steps), p.steps, b.mean);
    k_pack<<<blocks(b.fft_n), kBlock, 0, st>>>(b.raw, b.mean, p.steps, b.fft_n, b.packed, b.a);
}

// Radix sequence of the reference's stage loop (src/vulkan.cpp:234-241): radix 8 while
// at least three doublings remain (radix8), else radix 4 while two remain (radix4),
// else radix 2; the factor 3 of a 3*2^k length is the last stage, like the reference's
// final radix-3 combine.  Returns the buffer holding the spectrum.
float2* plain_fft(hipStream_t st, const PlainBuffers& b, bool radix8, bool radix4) {
    const uint32_t len = b.fft_n;
    const uint32_t pow2 = (len % 3 == 0) ? len / 3 : len;
    float2* in = b.a;
    float2* out = b.scratch;
    uint32_t ns = 1, done = 1;
    while (done < pow2) {
        const uint32_t left = pow2 / done;
        const uint32_t r = (radix8 && left >= 8) ? 8 : ((radix4 && left >= 4) ? 4 : 2);
        if (r == 8) k_stockham<8><<<blocks(len / 8), kBlock, 0, st>>>(in, out, b.tw_fft, len, ns);
        else if (r == 4) k_stockham<4><<<blocks(len / 4), kBlock, 0, st>>>(in, out, b.tw_fft, len, ns);
        else k_stockham<2><<<blocks(len / 2), kBlock, 0, st>>>(in, out, b.tw_fft, len, ns);
        ns *= r;
        done *= r;
        float2* t = in; in = out; out = t;
    }
    if (pow2 != len) {
        k_stockham<3><<<blocks(len / 3), kBlock, 0, st>>>(in, out, b.tw_fft, len, ns);
        float2* t = in; in = out; out = t;
    }
    return in;
}

void plain_spectrum(hipStream_t st, const PlainBuffers& b, const float2* z, bool fused_power, float* power) {
    const float inv_n = 1.0f / float(b.n);
    if (b.packed) {
        k_real<<<blocks(b.n / 2 + 1), kBlock, 0, st>>>(z, b.tw_real, b.n / 2, inv_n, b.X, fused_power ? power : nullptr);
    } else {
        const hipError_t e = hipMemcpyAsync(b.X, z, sizeof(float2) * (b.n / 2 + 1), hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) throw std::runtime_error(std::string("copy spectrum: ") + hipGetErrorString(e));
    }
}

void plain_power(hipStream_t st, const PlainBuffers& b, float* power) {
    k_power<<<blocks(b.n / 2), kBlock, 0, st>>>(b.X, b.n / 2, 1.0f / float(b.n), power);
}

void plain_harmonic(hipStream_t st, const float* power, float* summed, const brp_params& p, bool scatter,
                    uint32_t* slot) {
    const dim3 grid(blocks(p.fundamental), 5);
    if (scatter) {
        k_hs_init<<<grid, kBlock, 0, st>>>(power, summed, p);
        if (p.harmonic > p.window)
            k_hs_scatter<<<blocks(p.harmonic - p.window), kBlock, 0, st>>>(power, summed, p);
        k_hs_collect<<<grid, kBlock, 0, st>>>(summed, p, slot);
    } else {
        k_harmonic<<<grid, kBlock, 0, st>>>(power, summed, p, slot);
    }
}

}  // namespace brp_engine
