// oracle.cpp -- CPU restatement of the BRP4 per-template search path.
//
// TEST INFRASTRUCTURE ONLY: the checker and the CPU baseline.  The shipped library
// never links or calls it.  GPL-2.0-or-later, like the science it restates.
//
// Each function names the reference file:line it follows.  Parity status of each
// function (what pins it) is recorded in tests/test_oracle.py; nothing here is
// pinned against the reference's own binaries, which cannot be built in this image.
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_error;
int fail(const std::string& m) { g_error = m; return -1; }

constexpr double kPi = 3.14159265358979323846264338327950288;

// include/brp/lut.hpp:4-5 -- data, kept verbatim (the resampling indices depend on
// every bit).  Pinned against tests/golden/lut_table.json.
const float kSinTab[65] = {
    0.000000f, 0.098017f, 0.195090f, 0.290285f, 0.382683f, 0.471397f, 0.555570f, 0.634393f,
    0.707107f, 0.773010f, 0.831470f, 0.881921f, 0.923880f, 0.956940f, 0.980785f, 0.995185f,
    1.000000f, 0.995185f, 0.980785f, 0.956940f, 0.923880f, 0.881921f, 0.831470f, 0.773010f,
    0.707107f, 0.634393f, 0.555570f, 0.471397f, 0.382683f, 0.290285f, 0.195091f, 0.098017f,
    0.000000f, -0.098017f, -0.195090f, -0.290284f, -0.382683f, -0.471397f, -0.555570f, -0.634393f,
    -0.707107f, -0.773010f, -0.831469f, -0.881921f, -0.923880f, -0.956940f, -0.980785f, -0.995185f,
    -1.000000f, -0.995185f, -0.980785f, -0.956940f, -0.923880f, -0.881921f, -0.831470f, -0.773011f,
    -0.707107f, -0.634394f, -0.555570f, -0.471397f, -0.382684f, -0.290285f, -0.195091f, -0.098017f,
    -0.000000f};
const float kCosTab[65] = {
    1.000000f, 0.995185f, 0.980785f, 0.956940f, 0.923880f, 0.881921f, 0.831470f, 0.773010f,
    0.707107f, 0.634393f, 0.555570f, 0.471397f, 0.382683f, 0.290285f, 0.195090f, 0.098017f,
    0.000000f, -0.098017f, -0.195090f, -0.290285f, -0.382683f, -0.471397f, -0.555570f, -0.634393f,
    -0.707107f, -0.773010f, -0.831470f, -0.881921f, -0.923880f, -0.956940f, -0.980785f, -0.995185f,
    -1.000000f, -0.995185f, -0.980785f, -0.956940f, -0.923880f, -0.881921f, -0.831470f, -0.773011f,
    -0.707107f, -0.634393f, -0.555570f, -0.471397f, -0.382684f, -0.290285f, -0.195090f, -0.098017f,
    0.000000f, 0.098017f, 0.195090f, 0.290285f, 0.382683f, 0.471397f, 0.555570f, 0.634393f,
    0.707107f, 0.773010f, 0.831470f, 0.881921f, 0.923879f, 0.956940f, 0.980785f, 0.995185f,
    1.000000f};

bool pow2(uint32_t v) { return v && !(v & (v - 1)); }

// ---------------------------------------------------------------- FFT (Stockham) --
// Self-sorting mixed-radix Stockham, radices 4, 2 and 3.  Stage s with radix r and
// Ns = product of earlier radices reads v[q] = x[j + q*n/r], applies the twiddle
// e^{-2 pi i (j mod Ns) q / (Ns r)}, takes a length-r DFT and writes
// y[(j / Ns) * Ns * r + (j mod Ns) + q * Ns].  Output is in natural order.
template <class T>
struct Plan {
    size_t n = 0;
    std::vector<int> radix;
    std::vector<std::complex<T>> w;  // w[t] = e^{-2 pi i t / n}, computed in double
    explicit Plan(size_t len) : n(len) {
        size_t m = n;
        if (m % 3 == 0) { radix.push_back(3); m /= 3; }
        while (m % 4 == 0) { radix.push_back(4); m /= 4; }
        while (m % 2 == 0) { radix.push_back(2); m /= 2; }
        w.resize(n);
        for (size_t t = 0; t < n; ++t) {
            double a = -2.0 * kPi * double(t) / double(n);
            w[t] = {T(std::cos(a)), T(std::sin(a))};
        }
    }
    // Returns a pointer to whichever of a/b holds the result.
    std::complex<T>* run(std::complex<T>* a, std::complex<T>* b) const {
        using C = std::complex<T>;
        const T s3 = T(0.86602540378443864676372317075293618);
        size_t ns = 1;
        C* x = a; C* y = b;
        for (int r : radix) {
            size_t span = n / size_t(r);
            size_t tstride = n / (ns * size_t(r));
            for (size_t j = 0; j < span; ++j) {
                size_t k = j % ns;
                size_t out = (j / ns) * ns * size_t(r) + k;
                if (r == 2) {
                    C v0 = x[j], v1 = x[j + span] * w[k * tstride];
                    y[out] = v0 + v1; y[out + ns] = v0 - v1;
                } else if (r == 4) {
                    C v0 = x[j];
                    C v1 = x[j + span] * w[k * tstride];
                    C v2 = x[j + 2 * span] * w[2 * k * tstride];
                    C v3 = x[j + 3 * span] * w[3 * k * tstride];
                    C s02 = v0 + v2, d02 = v0 - v2, s13 = v1 + v3, d13 = v1 - v3;
                    C d13j(d13.imag(), -d13.real());  // -i * d13
                    y[out] = s02 + s13; y[out + ns] = d02 + d13j;
                    y[out + 2 * ns] = s02 - s13; y[out + 3 * ns] = d02 - d13j;
                } else {  // r == 3
                    C v0 = x[j];
                    C v1 = x[j + span] * w[k * tstride];
                    C v2 = x[j + 2 * span] * w[2 * k * tstride];
                    C s = v1 + v2, d = v1 - v2;
                    C c = v0 - T(0.5) * s;
                    C t(d.imag() * s3, -d.real() * s3);  // -i * sin(2pi/3) * d
                    y[out] = v0 + s; y[out + ns] = c + t; y[out + 2 * ns] = c - t;
                }
            }
            ns *= size_t(r);
            std::swap(x, y);
        }
        return x;
    }
};

// Real FFT of length n (even) through the half-length complex FFT and the split
// X[k] = 1/2 [(Z_k + conj Z_{M-k}) - i W_n^k (Z_k - conj Z_{M-k})], M = n / 2.
template <class T, class In>
void rfft(const In* x, size_t n, T* out) {
    using C = std::complex<T>;
    size_t m = n / 2;
    Plan<T> plan(m);
    std::vector<C> a(m), b(m);
    for (size_t i = 0; i < m; ++i) a[i] = {T(x[2 * i]), T(x[2 * i + 1])};
    C* z = plan.run(a.data(), b.data());
    for (size_t k = 0; k <= m; ++k) {
        C zk = z[k % m], zr = std::conj(z[(m - k) % m]);
        double ang = -2.0 * kPi * double(k) / double(n);
        C wk(T(std::cos(ang)), T(std::sin(ang)));
        C s = zk + zr, d = (zk - zr) * wk;
        C res = T(0.5) * (s + C(d.imag(), -d.real()));
        out[2 * k] = res.real(); out[2 * k + 1] = res.imag();
    }
}

// shaders/hs_sum.glsl:2-10 / shaders/harmonic.comp:5-13: dyadic FP32 sums in the
// historical order; index (i*k + 8) >> 4 is round-half-up of i*k/16.
inline void dyadic(const float* P, uint32_t i, float s[5]) {
    auto at = [&](uint32_t k) { return P[(i * k + 8) >> 4]; };
    float v = P[i];
    s[0] = v;
    v += at(8); s[1] = v;
    v += at(12) + at(4); s[2] = v;
    v += at(14) + at(10) + at(6) + at(2); s[3] = v;
    v += at(15) + at(13) + at(11) + at(9) + at(7) + at(5) + at(3) + at(1); s[4] = v;
}

}  // namespace

extern "C" {

const char* oracle_last_error(void) { return g_error.c_str(); }

void oracle_lut_table(float* out) {
    for (int i = 0; i < 65; ++i) { out[2 * i] = kSinTab[i]; out[2 * i + 1] = kCosTab[i]; }
}

// src/science.cpp:20-28
float oracle_lut_sin(float x) {
    float whole;
    float frac = std::modf((1.0f / 6.283185f) * x, &whole);
    if (frac < 0) frac += 1;
    int seg = int(frac * 64.0f + 0.5f);
    float d = 6.283185f * (frac - (1.0f / 64) * float(seg));
    float d2 = d;
    d2 *= 0.5f * d;
    return kSinTab[seg] + d * kCosTab[seg] - d2 * kSinTab[seg];
}

// src/science.cpp:29-35
float oracle_fused_sin(float x) {
    float whole;
    float frac = std::modf((1.0f / 6.283185f) * x, &whole);
    if (frac < 0) frac += 1;
    int seg = int(std::fma(frac, 64.0f, 0.5f));
    float d = 6.283185f * std::fma(-float(seg), 1.0f / 64, frac);
    float d2 = (0.5f * d) * d;
    return std::fma(-d2, kSinTab[seg], std::fma(d, kCosTab[seg], kSinTab[seg]));
}

// src/science.cpp:36-40
float oracle_sample_offset(uint32_t i, const oracle_params* p) {
    float t = float(i) * p->dt;
    if (p->flags & 2)
        return std::fma(p->tau * oracle_fused_sin(std::fma(p->omega, t, p->psi)), p->step_inv, -p->s0);
    return p->tau * oracle_lut_sin(p->omega * t + p->psi) * p->step_inv - p->s0;
}

// src/science.cpp:41-45
int oracle_resampling_steps(oracle_params* p) {
    uint32_t last = p->unpadded - 1;
    uint32_t s = last;
    while (s && float(s) - oracle_sample_offset(s, p) >= float(last)) --s;
    p->steps = s;
    if (!s) return fail("template produces empty series");
    return 0;
}

// src/science.cpp:46-65
int oracle_parameters(uint32_t n, uint32_t unpadded, float tau, float period, float psi, float dt,
                      oracle_params* out) {
    uint32_t radix = pow2(n) ? 1 : ((n % 3 == 0 && pow2(n / 3)) ? 3 : 0);
    if (!radix || n < 64 || n > 3u * (1u << 22) || unpadded < 32 || unpadded > n || dt <= 0 ||
        period <= 0 || tau < 0 || !std::isfinite(tau + period + psi + dt))
        return fail("invalid observation/template");
    oracle_params p;
    std::memset(&p, 0, sizeof p);
    p.n = n; p.unpadded = unpadded; p.fft_n = n; p.fft_radix = radix; p.fft_base = n / radix;
    uint32_t lg = 0; while (!((p.fft_base >> lg) & 1)) ++lg;
    p.logn = lg; p.tau = tau;
    p.omega = float(2.0 * kPi / double(period));
    p.psi = psi; p.dt = dt; p.step_inv = 1.0f / dt;
    p.s0 = float(double(tau) * std::sin(double(psi)) * double(p.step_inv));
    if (!std::isfinite(p.omega) || !std::isfinite(p.step_inv) || !std::isfinite(p.s0) ||
        !std::isfinite(p.tau * p.step_inv) || !std::isfinite(float(unpadded) * p.dt) ||
        !std::isfinite(p.omega * (float(unpadded) * p.dt) + p.psi))
        return fail("template exceeds finite FP32 arithmetic domain");
    if (oracle_resampling_steps(&p)) return -1;
    p.harmonic = n / 2; p.fundamental = n / 2; p.window = 8;
    const float thr[5] = {20, 30, 45, 70, 110};
    for (int h = 0; h < 5; ++h) p.threshold[h] = thr[h];
    p.capacity = 65536;
    *out = p;
    return 0;
}

// src/science.cpp:75-85 (sequential FP32 sum for the pad mean; +0.5 in double).
int oracle_resample(const float* input, size_t count, const oracle_params* p, float* out,
                    int32_t* index) {
    if (count != p->unpadded) return fail("input length mismatch");
    float sum = 0;
    for (uint32_t i = 0; i < p->steps; ++i) {
        int j = int(double(float(i) - oracle_sample_offset(i, p)) + 0.5);
        if (j < 0 || uint32_t(j) >= p->unpadded) return fail("resample index outside observation");
        out[i] = input[j];
        sum += out[i];
        if (index) index[i] = j;
    }
    float mean = sum / float(p->steps);
    for (uint32_t i = p->steps; i < p->n; ++i) out[i] = mean;
    return 0;
}

int oracle_rfft_f64(const double* x, size_t n, double* out) {
    if (n < 4 || (n & 1)) return fail("rfft length must be even");
    rfft<double>(x, n, out);
    return 0;
}

int oracle_rfft_f32(const float* x, size_t n, float* out) {
    if (n < 4 || (n & 1)) return fail("rfft length must be even");
    rfft<float>(x, n, out);
    return 0;
}

// src/science.cpp:114-118: P[k] = (re^2 + im^2) / n for 0 < k < n/2, P[0] = 0.  The
// spectrum here is the FP64 transform of the FP32 series (the reference uses FFTW
// single precision; the FP64 transform is the more accurate of the two).
int oracle_power(const float* resampled, size_t n, float* power, double* power64) {
    std::vector<double> spec(2 * (n / 2 + 1));
    rfft<double>(resampled, n, spec.data());
    for (size_t k = 0; k < n / 2; ++k) {
        double re = spec[2 * k], im = spec[2 * k + 1];
        double v = k ? (re * re + im * im) / double(n) : 0.0;
        power[k] = float(v);
        if (power64) power64[k] = v;
    }
    return 0;
}

// shaders/harmonic.comp:14-36 (gather form); level 0 copies the power.
int oracle_harmonic(const float* P, size_t count, const oracle_params* p, float* summed) {
    if (count < p->harmonic) return fail("power length mismatch");
    const uint32_t F = p->fundamental;
    for (uint32_t j = 0; j < F; ++j) summed[j] = P[j];
    for (uint32_t h = 1; h < 5; ++h) {
        const int64_t width = int64_t(1) << h;
        for (uint32_t j = 0; j < F; ++j) {
            int64_t begin = std::max<int64_t>(p->window, int64_t(j) * width - width / 2);
            int64_t end = std::min<int64_t>(p->harmonic, int64_t(j) * width + width / 2);
            float value = 0;
            if (begin < end) {
                float s[5];
                dyadic(P, uint32_t(begin), s);
                float first = s[h], best = first;
                for (int64_t i = begin + 1; i < end; ++i) {
                    dyadic(P, uint32_t(i), s);
                    best = std::max(best, s[h]);
                }
                value = best > p->threshold[h] ? best : first;
            }
            summed[size_t(h) * F + j] = value;
        }
    }
    return 0;
}

// shaders/hs_init.comp:6-18, hs_scatter.comp:7-16 (scatter form).
int oracle_harmonic_scatter(const float* P, size_t count, const oracle_params* p, float* summed) {
    if (count < p->harmonic) return fail("power length mismatch");
    const uint32_t F = p->fundamental;
    for (uint32_t j = 0; j < F; ++j) summed[j] = P[j];
    for (uint32_t h = 1; h < 5; ++h) {
        const int64_t width = int64_t(1) << h;
        for (uint32_t j = 0; j < F; ++j) {
            int64_t begin = std::max<int64_t>(p->window, int64_t(j) * width - width / 2);
            int64_t end = std::min<int64_t>(p->harmonic, int64_t(j) * width + width / 2);
            float s[5] = {0, 0, 0, 0, 0};
            if (begin < end) dyadic(P, uint32_t(begin), s);
            summed[size_t(h) * F + j] = begin < end ? s[h] : 0.0f;
        }
    }
    for (uint32_t i = p->window; i < p->harmonic; ++i) {
        float s[5];
        dyadic(P, i, s);
        for (uint32_t h = 1; h < 5; ++h) {
            uint32_t j = (i + (1u << (h - 1))) >> h;
            float& cell = summed[size_t(h) * F + j];
            if (j < F && s[h] > p->threshold[h]) cell = std::max(cell, s[h]);
        }
    }
    return 0;
}

// src/science.cpp:106-112
size_t oracle_select(const float* summed, const oracle_params* p, oracle_candidate* out, size_t cap) {
    size_t count = 0;
    for (uint32_t h = 0; h < 5; ++h)
        for (uint32_t j = p->window; j < p->fundamental; ++j) {
            float x = summed[size_t(h) * p->fundamental + j];
            if (x > p->threshold[h]) {
                if (count < cap) out[count] = {j, h, x, 0};
                ++count;
            }
        }
    return count;  // generated in (level, bin) order already
}

// src/science.cpp:113-120
int oracle_reference(const float* input, size_t count, const oracle_params* p, float* resampled,
                     float* power, float* summed) {
    if (oracle_resample(input, count, p, resampled, nullptr)) return -1;
    if (oracle_power(resampled, p->n, power, nullptr)) return -1;
    return oracle_harmonic(power, p->n / 2, p, summed);
}

// ------------------------------------------------------------- bank semantics --
// Q(x; 2k) = e^{-x/2} sum_{i<k} (x/2)^i / i!  (upper tail of chi^2 with even dof),
// evaluated in the log domain so that large x does not underflow early.
static double log_chisq_Q_even(double x, uint32_t k) {
    if (x <= 0) return 0.0;
    double h = 0.5 * x, lh = std::log(h);
    double m = -1e300;
    std::vector<double> terms(k);
    double lf = 0;  // log(i!)
    for (uint32_t i = 0; i < k; ++i) {
        if (i) lf += std::log(double(i));
        terms[i] = double(i) * lh - lf;
        m = std::max(m, terms[i]);
    }
    double s = 0;
    for (double t : terms) s += std::exp(t - m);
    return -h + m + std::log(s);
}

double oracle_chisq_Q(double x, double dof) {
    uint32_t k = uint32_t(dof / 2);
    double lq = log_chisq_Q_even(x, k);
    return lq < -745.0 ? 0.0 : std::exp(lq);
}

// Inverse of Q by bisection on log Q (monotone decreasing in x), to full double precision.
double oracle_chisq_Qinv(double q, double dof) {
    uint32_t k = uint32_t(dof / 2);
    if (!(q > 0) || !(q < 1)) return std::nan("");
    double target = std::log(q);
    double lo = 0, hi = 1;
    while (log_chisq_Q_even(hi, k) > target) hi *= 2;
    for (int it = 0; it < 200; ++it) {
        double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (log_chisq_Q_even(mid, k) > target) lo = mid; else hi = mid;
    }
    return 0.5 * (lo + hi);
}

// src/workload.cpp:108-125
int oracle_search_parameters(uint32_t n, uint32_t samples, double sample_us, double period,
                             double tau, double psi, double f0, double false_alarm, uint32_t window,
                             oracle_params* out) {
    if (!std::isfinite(f0 + false_alarm) || f0 <= 0 || false_alarm <= 0 || false_alarm >= 1 || !window)
        return fail("invalid search parameters");
    oracle_params p;
    if (oracle_parameters(n, samples, float(tau), float(period), float(psi), float(sample_us * 1e-6), &p))
        return -1;
    p.flags = 2 | 4;
    p.s0 = (p.tau * std::sin(p.psi)) * p.step_inv;  // float sin, as the reference
    if (oracle_resampling_steps(&p)) return -1;
    double seconds = double(n) * sample_us * 1e-6;
    p.window = uint32_t(window * 0.5 + 0.5);
    if (p.window >= p.n / 2 + 1) return fail("search window exceeds spectrum");
    double limit = double(p.n / 2 + 1 - p.window);
    p.fundamental = uint32_t(std::min(limit, f0 * seconds + 0.5));
    p.harmonic = uint32_t(std::min(limit, 16 * f0 * seconds + 0.5));
    float probability = float(1.0 - std::pow(1.0 - false_alarm, 1.0 / double(p.n / 2 + 1)));
    if (!(probability > 0 && probability < 1)) return fail("false-alarm probability outside domain");
    for (uint32_t h = 0; h < 5; ++h)
        p.threshold[h] = float(0.5 * oracle_chisq_Qinv(double(probability), 2.0 * double(1u << h)));
    if (p.fundamental <= p.window) return fail("empty frequency search");
    *out = p;
    return 0;
}

// src/workload.cpp:126
void oracle_toplist_thresholds(const oracle_bank_row* st, uint32_t rows, oracle_params* p) {
    for (uint32_t h = 0; h < 5; ++h)
        p->threshold[h] = std::max(p->threshold[h], float(st[(h + 1) * rows - 1].power));
}

static bool power_desc(const oracle_bank_row& a, const oracle_bank_row& b) { return a.power > b.power; }
static bool sig_desc(const oracle_bank_row& a, const oracle_bank_row& b) {
    return a.significance > b.significance;
}

// src/workload.cpp:129-141 (the reference's glibc qsort is a stable merge sort here).
int oracle_toplist_retain(oracle_bank_row* st, uint32_t rows, const oracle_candidate* cand,
                          size_t count, double period, double tau, double psi, const oracle_params* p) {
    for (size_t n = 0; n < count; ++n) {
        const oracle_candidate c = cand[n];
        if (c.level >= 5 || c.bin < p->window || c.bin >= p->fundamental || !std::isfinite(c.power))
            return fail("invalid compact candidate");
        uint32_t first = c.level * rows, last = first + rows - 1;
        if (c.power <= p->threshold[c.level] || c.power <= st[last].power) continue;
        int64_t place = last;
        for (uint32_t i = first; i <= last; ++i)
            if (st[i].bin == c.bin) { place = st[i].power < c.power ? int64_t(i) : -1; break; }
        if (place >= int64_t(first)) {
            st[place] = {double(c.power), double(float(period)), double(float(tau)), double(float(psi)),
                         0.0, 1u << c.level, c.bin};
            std::stable_sort(st + first, st + first + rows, power_desc);
        }
    }
    return 0;
}

// src/workload.cpp:142-159
size_t oracle_toplist_final(const oracle_bank_row* st, uint32_t rows, oracle_bank_row* out, size_t cap) {
    std::vector<oracle_bank_row> copy(st, st + 5 * size_t(rows));
    for (auto& c : copy) {
        if (!c.harmonics) { c.significance = -10; continue; }
        double q = oracle_chisq_Q(2 * c.power, 2.0 * c.harmonics);
        c.significance = q > 0 ? -std::log10(q) : 320;
        c.power /= double(float(std::sqrt(float(c.harmonics))));
    }
    std::stable_sort(copy.begin(), copy.end(), sig_desc);
    size_t n = 0;
    while (n < rows && copy[0].significance > 0) {
        if (n < cap) out[n] = copy[0];
        ++n;
        uint32_t bin = copy[0].bin;
        for (auto& c : copy) if (c.bin == bin) c.significance = -10;
        std::stable_sort(copy.begin(), copy.end(), sig_desc);
    }
    return n;
}

// ---------------------------------------------------------------- CPU baseline --
// One template: resample -> FP32 real FFT -> FP32 power -> harmonic -> select, the
// work of src/reference_adapter.cpp:43-50 with this file's FP32 FFT in place of FFTW.
static void cpu_template(const float* input, size_t count, const oracle_params& p,
                         std::vector<float>& rs, std::vector<float>& spec, std::vector<float>& pw,
                         std::vector<float>& sum, std::vector<oracle_candidate>& cand) {
    oracle_resample(input, count, &p, rs.data(), nullptr);
    rfft<float>(rs.data(), p.n, spec.data());
    const float inv = 1.0f / float(p.n);
    pw[0] = 0;
    for (uint32_t k = 1; k < p.n / 2; ++k)
        pw[k] = float(double(inv) * (double(spec[2 * k]) * spec[2 * k] + double(spec[2 * k + 1]) * spec[2 * k + 1]));
    oracle_harmonic(pw.data(), pw.size(), &p, sum.data());
    oracle_select(sum.data(), &p, cand.data(), cand.size());
}

// src/workload.cpp:29-95 (decode and preprocess), with the running median and RFI rule
// stated in oracle.h (the historical rngmed and the GSL zap noise are not available).
int oracle_preprocess(const uint8_t* bytes, size_t samples, uint32_t bits, double scale, uint32_t n,
                      uint32_t window, float rfi_ratio, float* out, float* decoded, uint32_t* rfi_bins) {
    return oracle_preprocess_zap(bytes, samples, bits, scale, n, window, rfi_ratio, nullptr, 0, 0.0, 1.0, out, decoded,
                                 rfi_bins);
}

// SplitMix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number generators",
// OOPSLA 2014) finaliser, and the counter-based uniform in [0, 1) the library defines
// over it: mix(mix(seed + golden gamma) + counter * 0xD1B54A32D192ED03) >> 11, times 2^-53.
static uint64_t splitmix_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double oracle_uniform(uint64_t seed, uint64_t counter) {
    const uint64_t z = splitmix_mix(splitmix_mix(seed + 0x9E3779B97F4A7C15ull) + counter * 0xD1B54A32D192ED03ull);
    return double(z >> 11) * 0x1.0p-53;
}

int oracle_preprocess_zap(const uint8_t* bytes, size_t samples, uint32_t bits, double scale, uint32_t n,
                          uint32_t window, float rfi_ratio, const double* zap, size_t zap_ranges, double sample_us,
                          double padding, float* out, float* decoded, uint32_t* rfi_bins) {
    if (!bytes || !out || (bits != 4 && bits != 8) || samples > n || n % 2 || !window || (window & 1) ||
        window > n / 2 || (zap_ranges && !zap))
        return fail("invalid preprocessing arguments");
    std::vector<double> time(n, 0.0);
    for (size_t i = 0; i < samples; ++i) {
        double v;
        if (bits == 4) v = double((i & 1) ? (bytes[i / 2] & 15) : (bytes[i / 2] >> 4));
        else v = double(int8_t(bytes[i]));
        const float f = float(v / scale);  // src/workload.cpp:52-53
        time[i] = f;
        if (decoded) decoded[i] = f;
    }
    const size_t m = n / 2, bins = m + 1;
    std::vector<double> spec(2 * bins);
    rfft<double>(time.data(), n, spec.data());
    std::vector<float> power(bins, 0.0f);
    for (size_t k = 1; k < bins; ++k) power[k] = float(spec[2 * k] * spec[2 * k] + spec[2 * k + 1] * spec[2 * k + 1]);
    const size_t nmed = bins - window + 1;
    std::vector<float> median(nmed), win(window);
    for (size_t i = 0; i < nmed; ++i) {
        std::copy(power.begin() + i, power.begin() + i + window, win.begin());
        std::nth_element(win.begin(), win.begin() + window / 2, win.end());
        const float hi = win[window / 2];
        const float lo = *std::max_element(win.begin(), win.begin() + window / 2);
        median[i] = float((double(lo) + double(hi)) * 0.5);
        if (!(median[i] > 0) || !std::isfinite(median[i])) return fail("invalid whitening median");
    }
    uint32_t rfi = 0;
    std::vector<std::complex<double>> full(n, {0.0, 0.0});
    for (size_t k = window / 2; k < bins - window / 2; ++k) {
        const float med = median[k - window / 2];
        std::complex<double> x(spec[2 * k], spec[2 * k + 1]);
        if (rfi_ratio > 0 && power[k] > rfi_ratio * med) {
            x *= std::sqrt(double(med) / double(power[k]));
            ++rfi;
        }
        x *= double(float(std::sqrt(0.69314718055994530942 / double(med))));
        full[k] = x;
        if (k < m) full[n - k] = std::conj(x);
    }
    // Zap list (src/workload.cpp:76-86): after whitening, each listed range's bins
    // uint32(f T + 0.5) .. (inclusive) get Gaussian noise of sigma sqrt(2)/2 sqrt(padding)
    // per component; the band edges stay zero.  Noise: Box-Muller over oracle_uniform
    // (counters 2k, 2k + 1) seeded by the bit pattern of the first decoded sample.
    if (zap_ranges) {
        const double seconds = double(n) * sample_us * 1e-6;
        const float first_sample = bits == 4 ? float(double(bytes[0] >> 4) / scale)
                                             : float(double(int8_t(bytes[0])) / scale);
        uint32_t seed32;
        std::memcpy(&seed32, &first_sample, 4);
        const double sigma = std::sqrt(2.0) / 2 * std::sqrt(padding);
        for (size_t r = 0; r < zap_ranges; ++r) {
            const double lower = zap[2 * r], upper = zap[2 * r + 1];
            if (!std::isfinite(lower + upper) || lower < 0 || upper < lower) return fail("invalid zap range");
            const double first = lower * seconds + 0.5, last = upper * seconds + 0.5;
            if (last >= double(bins)) return fail("zap exceeds Nyquist");
            for (uint32_t k = uint32_t(first); k <= uint32_t(last); ++k) {
                if (k < window / 2 || k >= bins - window / 2) continue;
                const double u1 = oracle_uniform(seed32, 2ull * k), u2 = oracle_uniform(seed32, 2ull * k + 1);
                const double radius = std::sqrt(-2.0 * std::log(1.0 - u1));
                const double angle = 6.283185307179586476925286766559 * u2;
                const std::complex<double> x(float(sigma * radius * std::cos(angle)),
                                             float(sigma * radius * std::sin(angle)));
                full[k] = x;
                if (k < m) full[n - k] = std::conj(x);
            }
        }
    }
    // Unnormalised inverse over the full Hermitian spectrum: y_j = Re FFT(conj X)_j.
    for (auto& c : full) c = std::conj(c);
    Plan<double> plan(n);
    std::vector<std::complex<double>> scratch(n);
    const std::complex<double>* y = plan.run(full.data(), scratch.data());
    const double inv = 1.0 / std::sqrt(double(n));
    for (size_t j = 0; j < samples; ++j) out[j] = float(y[j].real() * inv);
    if (rfi_bins) *rfi_bins = rfi;
    return 0;
}

int oracle_bank_candidates(const float* input, size_t count, const oracle_params* tpl, size_t ntpl,
                           uint32_t threads, uint32_t fft32, oracle_candidate* out, size_t cap, uint32_t* counts) {
    if (!threads) threads = 1;
    std::vector<std::string> errors(threads);
    std::vector<std::thread> pool;
    for (uint32_t w = 0; w < threads; ++w)
        pool.emplace_back([&, w] {
            std::vector<float> rs, spec, pw, sum;
            std::vector<oracle_candidate> tmp(cap + 1);
            for (size_t t = w; t < ntpl; t += threads) {
                const oracle_params& p = tpl[t];
                rs.resize(p.n); spec.resize(p.n + 2); pw.resize(p.n / 2); sum.resize(5 * size_t(p.fundamental));
                size_t k;
                if (fft32) {
                    cpu_template(input, count, p, rs, spec, pw, sum, tmp);
                    k = oracle_select(sum.data(), &p, out + t * cap, cap);
                } else {
                    if (oracle_reference(input, count, &p, rs.data(), pw.data(), sum.data())) {
                        errors[w] = g_error;
                        return;
                    }
                    k = oracle_select(sum.data(), &p, out + t * cap, cap);
                }
                if (k > cap) {
                    errors[w] = "template " + std::to_string(t) + " has " + std::to_string(k) + " candidates (cap " +
                                std::to_string(cap) + ")";
                    return;
                }
                counts[t] = uint32_t(k);
            }
        });
    for (auto& th : pool) th.join();
    for (const auto& e : errors)
        if (!e.empty()) return fail(e);
    return 0;
}

double oracle_benchmark_cpu(const float* input, size_t count, const oracle_params* tpl,
                            size_t ntpl, uint32_t threads, uint32_t reps) {
    if (!threads) threads = 1;
    if (!reps) reps = 1;
    std::vector<double> walls;
    for (uint32_t r = 0; r <= reps; ++r) {
        auto t0 = std::chrono::steady_clock::now();
        std::vector<std::thread> pool;
        for (uint32_t w = 0; w < threads; ++w)
            pool.emplace_back([&, w] {
                std::vector<float> rs, spec, pw, sum;
                std::vector<oracle_candidate> cand(65536);
                for (size_t t = w; t < ntpl; t += threads) {
                    const oracle_params& p = tpl[t];
                    rs.resize(p.n); spec.resize(p.n + 2); pw.resize(p.n / 2);
                    sum.resize(5 * size_t(p.fundamental));
                    cpu_template(input, count, p, rs, spec, pw, sum, cand);
                }
            });
        for (auto& th : pool) th.join();
        double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (r) walls.push_back(ms);
    }
    std::sort(walls.begin(), walls.end());
    return walls[walls.size() / 2];
}

}  // extern "C"
