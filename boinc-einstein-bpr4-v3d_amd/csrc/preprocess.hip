// preprocess.hip -- observation preprocessing on the device: decode of the DD sample
// block and spectral whitening by a running median, with RFI bins replaced by that
// median.  It takes the place of the reference's load_observation / preprocess
// (src/workload.cpp:29-95) ahead of the per-template search:
//
//   decode     y[2i] = float(double(b >> 4) / scale), y[2i+1] = float(double(b & 15) / scale)
//              (4-bit, high nibble first, src/workload.cpp:50-53) or
//              float(double(int8(b)) / scale) (8-bit)
//   spectrum   X = rfft_n(y zero-padded to n); P[k] = float(re^2 + im^2) (in double), P[0] = 0
//   median     m[i] = median of P[i .. i + w - 1] for i <= n/2 + 1 - w (even w: the mean of
//              the two middle values, a definition of this build: the historical rngmed
//              convention is not available)
//   whiten     X[k] *= float(sqrt(ln 2 / m[k - w/2])) for w/2 <= k < n/2 + 1 - w/2, all
//              other bins zero (src/workload.cpp:66-70, 86)
//   RFI        new behaviour (the reference zaps a given list with GSL noise instead,
//              src/workload.cpp:76-85): a bin with P[k] > ratio * m[k - w/2] is first
//              scaled to |X[k]|^2 = m[k - w/2], keeping its phase
//   zap        after whitening, the bins of each listed range [lower, upper] Hz,
//              uint32(f * T + 0.5) with T = n * sample_us * 1e-6, inclusive, become
//              sigma * (g1, g2) with sigma = sqrt(2)/2 * sqrt(padding)
//              (src/workload.cpp:76-85); bins outside [w/2, n/2 + 1 - w/2) stay zero
//              (src/workload.cpp:86).  g1, g2: Box-Muller from the counter-based uniform
//              of this library (brp_uniform, counter 2k and 2k + 1 for bin k), seeded by
//              the bit pattern of the first decoded sample as the reference seeds its
//              generator; the reference draws from GSL's taus2 ziggurat instead, which
//              is not available, so the zapped values are this build's own.
//   inverse    out[j] = c2r_n(X)[j] / sqrt(n) for j < samples (src/workload.cpp:87-92)
//
// The transforms are the plain engine's Stockham passes: the forward one packs pairs
// into n/2 complex points and splits; the inverse undoes the split,
// E = (X_k + conj X_{M-k}) / 2, O = W_n^{-k} (X_k - conj X_{M-k}) / 2, Z = E + i O, and
// takes conj(FFT(conj Z)) / M.
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "device.h"
#include "engine.h"

using namespace brp_dev;

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kMedianOut = 1024;   // medians per block
constexpr uint32_t kMedianSort = 4096;  // sorted entries per block: kMedianOut + w - 1 <= 4096

uint32_t blocks(size_t n) { return uint32_t((n + kBlock - 1) / kBlock); }

__global__ __launch_bounds__(kBlock) void k_decode(const uint8_t* __restrict__ bytes, uint32_t samples,
                                                   uint32_t bits, double scale, uint32_t m,
                                                   float* __restrict__ decoded, float2* __restrict__ a) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;  // output pair
    if (p >= m) return;
    float v[2] = {0.0f, 0.0f};
    for (int h = 0; h < 2; ++h) {
        const uint32_t i = 2 * p + h;
        if (i >= samples) continue;
        if (bits == 4) {
            const uint32_t b = bytes[i / 2];
            v[h] = float(double((i & 1u) ? (b & 15u) : (b >> 4)) / scale);
        } else {
            v[h] = float(double(int8_t(bytes[i])) / scale);
        }
        if (decoded) decoded[i] = v[h];
    }
    a[p] = make_float2(v[0], v[1]);
}

// P[k] = float(re^2 + im^2) with the products in double (src/workload.cpp:64), P[0] = 0.
__global__ __launch_bounds__(kBlock) void k_spectrum_power(const float2* __restrict__ X, uint32_t bins,
                                                           float* __restrict__ P) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= bins) return;
    const double re = X[k].x, im = X[k].y;
    P[k] = k == 0 ? 0.0f : float(re * re + im * im);
}

// Running median of w (even, <= kMedianSort - kMedianOut + 1) values: block b owns the
// medians o0 .. o0 + kMedianOut - 1; it sorts the values of their union window (value
// bits above position, so the sort is by value, ties by position) in LDS with a bitonic
// network, then each thread walks the sorted list counting the entries inside its own
// window up to the two middle ranks.  Powers are >= 0, so float bits order like values.
__global__ __launch_bounds__(1024) void k_running_median(const float* __restrict__ P, uint32_t bins, uint32_t w,
                                                         float* __restrict__ median, uint32_t* __restrict__ bad) {
    __shared__ uint64_t key[kMedianSort];
    const uint32_t t = threadIdx.x;
    const uint32_t o0 = blockIdx.x * kMedianOut;
    const uint32_t nmed = bins - w + 1;
    const uint32_t span = min(kMedianOut + w - 1, bins - o0);
    for (uint32_t e = t; e < kMedianSort; e += 1024)
        key[e] = e < span ? (uint64_t(__float_as_uint(P[o0 + e])) << 32) | e : ~0ull;
    __syncthreads();
    for (uint32_t size = 2; size <= kMedianSort; size <<= 1)
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t e = t; e < kMedianSort / 2; e += 1024) {
                const uint32_t lo = 2 * e - (e & (stride - 1));
                const uint32_t hi = lo + stride;
                const bool up = (lo & size) == 0;
                const uint64_t a = key[lo], b = key[hi];
                if ((a > b) == up) { key[lo] = b; key[hi] = a; }
            }
            __syncthreads();
        }
    const uint32_t o = o0 + t;
    if (t >= kMedianOut || o >= nmed) return;
    const uint32_t first = t, last = t + w;  // positions of this window within the block
    const uint32_t r1 = w / 2 - 1, r2 = w / 2;
    uint32_t seen = 0;
    float a = 0.0f, b = 0.0f;
    for (uint32_t e = 0; e < kMedianSort; ++e) {  // the window's w entries are all inside the sort
        const uint64_t kv = key[e];
        const uint32_t pos = uint32_t(kv);
        if (pos < first || pos >= last) continue;
        if (seen == r1) a = __uint_as_float(uint32_t(kv >> 32));
        if (seen == r2) { b = __uint_as_float(uint32_t(kv >> 32)); break; }
        ++seen;
    }
    const float m = float((double(a) + double(b)) * 0.5);
    median[o] = m;
    if (!(m > 0.0f) || !isfinite(m)) atomicOr(bad, 1u);
}

// Whitening and RFI replacement in place on X[0 .. bins - 1]; counts replaced bins.
__global__ __launch_bounds__(kBlock) void k_whiten(float2* __restrict__ X, const float* __restrict__ P,
                                                   const float* __restrict__ median, uint32_t bins, uint32_t w,
                                                   float rfi_ratio, uint32_t* __restrict__ rfi) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= bins) return;
    if (k < w / 2 || k >= bins - w / 2) {
        X[k] = make_float2(0.0f, 0.0f);
        return;
    }
    const float m = median[k - w / 2];
    float2 x = X[k];
    if (rfi_ratio > 0.0f && P[k] > rfi_ratio * m) {
        const float s = float(sqrt(double(m) / double(P[k])));
        x = make_float2(x.x * s, x.y * s);
        atomicAdd(rfi, 1u);
    }
    const float factor = float(sqrt(0.69314718055994530942 / double(m)));
    X[k] = make_float2(x.x * factor, x.y * factor);
}

// Zapped bins (bin, value) in place, for bins inside the whitened band.
__global__ __launch_bounds__(kBlock) void k_zap(float2* __restrict__ X, uint32_t bins, uint32_t w,
                                                const uint32_t* __restrict__ bin, const float2* __restrict__ value,
                                                uint32_t count) {
    const uint32_t z = blockIdx.x * kBlock + threadIdx.x;
    if (z >= count) return;
    const uint32_t k = bin[z];
    if (k >= w / 2 && k < bins - w / 2) X[k] = value[z];
}

// conj(Z[k]) for the inverse of the real split, k < M (X[0] and X[M] are zero after
// whitening, and the general formula covers them anyway).
__global__ __launch_bounds__(kBlock) void k_inverse_split(const float2* __restrict__ X, const float2* __restrict__ twr,
                                                          uint32_t m, float2* __restrict__ a) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= m) return;
    const cf xk = load(X + k), xr = conj(load(X + (m - k)));
    const cf e = scale(add(xk, xr), 0.5f);
    const cf o = scale(mul(sub(xk, xr), conj(load(twr + k))), 0.5f);
    const cf z = cf{e.x - o.y, e.y + o.x};  // E + i O
    store(a + k, conj(z));
}

__global__ __launch_bounds__(kBlock) void k_unpack(const float2* __restrict__ w, uint32_t m, uint32_t samples,
                                                   float s, float* __restrict__ out) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= m) return;
    const float2 v = w[p];
    if (2 * p < samples) out[2 * p] = v.x * s;
    if (2 * p + 1 < samples) out[2 * p + 1] = -v.y * s;
}

std::vector<float2> twiddle_table(uint64_t len, uint64_t count) {
    std::vector<float2> tw(count);
    for (uint64_t t = 0; t < count; ++t) {
        const double ang = -6.283185307179586476925286766559 * double(t % len) / double(len);
        tw[t] = make_float2(float(std::cos(ang)), float(std::sin(ang)));
    }
    return tw;
}

struct DeviceBuffers {
    std::vector<void*> held;
    template <class T>
    T* get(size_t count) {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess)
            throw std::runtime_error("preprocess: device allocation failed");
        held.push_back(p);
        return static_cast<T*>(p);
    }
    ~DeviceBuffers() {
        for (void* p : held) hipFree(p);
    }
};

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

namespace brp_engine {

void preprocess_observation(const uint8_t* bytes, size_t samples, uint32_t bits, double scale, uint32_t n,
                            uint32_t window, float rfi_ratio, const double* zap, size_t zap_ranges, double sample_us,
                            double padding, float* out, float* decoded, uint32_t* rfi_bins) {
    const uint32_t m = n / 2, bins = m + 1;
    // Zap list first (src/workload.cpp:76-85), so that a bad range fails before any
    // device work: bins and their noise values, one entry per bin (a bin named by two
    // ranges gets the same value either way).
    std::vector<uint32_t> zap_bin;
    std::vector<float2> zap_value;
    if (zap_ranges) {
        const double seconds = double(n) * sample_us * 1e-6;
        const float first_sample = bits == 4 ? float(double(bytes[0] >> 4) / scale)
                                             : float(double(int8_t(bytes[0])) / scale);
        uint32_t seed32;
        std::memcpy(&seed32, &first_sample, 4);
        const double sigma = std::sqrt(2.0) / 2 * std::sqrt(padding);
        std::vector<uint8_t> seen(bins, 0);
        for (size_t r = 0; r < zap_ranges; ++r) {
            const double lower = zap[2 * r], upper = zap[2 * r + 1];
            if (!std::isfinite(lower + upper) || lower < 0 || upper < lower) throw std::runtime_error("invalid zap range");
            const double first = lower * seconds + 0.5, last = upper * seconds + 0.5;
            if (last >= double(bins)) throw std::runtime_error("zap exceeds Nyquist");
            for (uint32_t k = uint32_t(first); k <= uint32_t(last); ++k) {
                if (seen[k]) continue;
                seen[k] = 1;
                const double u1 = brp_uniform(seed32, 2ull * k), u2 = brp_uniform(seed32, 2ull * k + 1);
                const double radius = std::sqrt(-2.0 * std::log(1.0 - u1));
                const double angle = 6.283185307179586476925286766559 * u2;
                zap_bin.push_back(k);
                zap_value.push_back(make_float2(float(sigma * radius * std::cos(angle)),
                                                float(sigma * radius * std::sin(angle))));
            }
        }
    }
    DeviceBuffers d;
    hipStream_t st = nullptr;
    check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { hipStreamSynchronize(s); hipStreamDestroy(s); }
    } guard{st};
    const size_t nbytes = bits == 4 ? samples / 2 : samples;
    uint8_t* dbytes = d.get<uint8_t>(nbytes);
    check(hipMemcpyAsync(dbytes, bytes, nbytes, hipMemcpyHostToDevice, st), "copy samples");
    PlainBuffers b;
    b.n = n;
    b.fft_n = m;
    b.packed = 1;
    b.a = d.get<float2>(m + 1);
    b.scratch = d.get<float2>(m + 1);
    b.X = d.get<float2>(bins);
    const std::vector<float2> tw_fft = twiddle_table(m, m), tw_real = twiddle_table(n, bins);
    float2* dtw_fft = d.get<float2>(m);
    float2* dtw_real = d.get<float2>(bins);
    check(hipMemcpyAsync(dtw_fft, tw_fft.data(), sizeof(float2) * m, hipMemcpyHostToDevice, st), "copy twiddles");
    check(hipMemcpyAsync(dtw_real, tw_real.data(), sizeof(float2) * bins, hipMemcpyHostToDevice, st), "copy twiddles");
    b.tw_fft = dtw_fft;
    b.tw_real = dtw_real;
    float* ddecoded = decoded ? d.get<float>(samples) : nullptr;
    float* P = d.get<float>(bins);
    const uint32_t nmed = bins - window + 1;
    float* med = d.get<float>(nmed);
    uint32_t* counters = d.get<uint32_t>(2);  // {invalid median, RFI bins}
    float* dout = d.get<float>(samples);
    check(hipMemsetAsync(counters, 0, 2 * sizeof(uint32_t), st), "clear counters");

    k_decode<<<blocks(m), kBlock, 0, st>>>(dbytes, uint32_t(samples), bits, scale, m, ddecoded, b.a);
    const float2* z = plain_fft(st, b, true, true);
    plain_spectrum(st, b, z, false, nullptr);
    k_spectrum_power<<<blocks(bins), kBlock, 0, st>>>(b.X, bins, P);
    k_running_median<<<(nmed + kMedianOut - 1) / kMedianOut, 1024, 0, st>>>(P, bins, window, med, counters);
    k_whiten<<<blocks(bins), kBlock, 0, st>>>(b.X, P, med, bins, window, rfi_ratio, counters + 1);
    if (!zap_bin.empty()) {
        const uint32_t nz = uint32_t(zap_bin.size());
        uint32_t* dbin = d.get<uint32_t>(nz);
        float2* dval = d.get<float2>(nz);
        check(hipMemcpyAsync(dbin, zap_bin.data(), sizeof(uint32_t) * nz, hipMemcpyHostToDevice, st), "copy zap bins");
        check(hipMemcpyAsync(dval, zap_value.data(), sizeof(float2) * nz, hipMemcpyHostToDevice, st), "copy zap values");
        k_zap<<<blocks(nz), kBlock, 0, st>>>(b.X, bins, window, dbin, dval, nz);
    }
    k_inverse_split<<<blocks(m), kBlock, 0, st>>>(b.X, b.tw_real, m, b.a);
    const float2* wv = plain_fft(st, b, true, true);
    k_unpack<<<blocks(m), kBlock, 0, st>>>(wv, m, uint32_t(samples), float(std::sqrt(double(n)) / double(m)), dout);
    check(hipGetLastError(), "preprocess launch");
    uint32_t host_counters[2] = {0, 0};
    check(hipMemcpyAsync(host_counters, counters, sizeof(host_counters), hipMemcpyDeviceToHost, st), "read counters");
    check(hipMemcpyAsync(out, dout, sizeof(float) * samples, hipMemcpyDeviceToHost, st), "read series");
    if (decoded) check(hipMemcpyAsync(decoded, ddecoded, sizeof(float) * samples, hipMemcpyDeviceToHost, st), "read decoded");
    check(hipStreamSynchronize(st), "preprocess");
    if (host_counters[0]) throw std::runtime_error("invalid whitening median");
    if (rfi_bins) *rfi_bins = host_counters[1];
}

}  // namespace brp_engine
