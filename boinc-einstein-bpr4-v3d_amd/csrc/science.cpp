// science.cpp -- host-side science of the BRP4 per-template path: template parameters
// (bench and bank semantics), the chi^2 thresholds and significances, the bank toplist,
// the results file and the seeded synthetic workloads.
//
// Compiled with -ffp-contract=off: the FP32 expressions below must round exactly as
// written, because resampling_steps() decides the resampled length with them
// (reference CMakeLists.txt:14).  GPL-2.0-or-later.
#include "brp_hip.h"
#include "errors.h"
#include "lut.h"
#include "toplist_state.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

bool is_pow2(uint32_t v) { return v && !(v & (v - 1)); }

// FP32 LUT sine with separately rounded operations (src/science.cpp:20-28).
float lut_sin(float phase) {
    float whole;
    float frac = std::modf((1.0f / 6.283185f) * phase, &whole);
    if (frac < 0) frac += 1;
    const int seg = int(frac * 64.0f + 0.5f);
    const float d = 6.283185f * (frac - (1.0f / 64) * float(seg));
    float d2 = d;
    d2 *= 0.5f * d;
    const float s = brp_lut::kTable[2 * seg], c = brp_lut::kTable[2 * seg + 1];
    return s + d * c - d2 * s;
}

// Nested-FMA variant used with flag 2 (src/science.cpp:29-35).
float fma_sin(float phase) {
    float whole;
    float frac = std::modf((1.0f / 6.283185f) * phase, &whole);
    if (frac < 0) frac += 1;
    const int seg = int(std::fma(frac, 64.0f, 0.5f));
    const float d = 6.283185f * std::fma(-float(seg), 1.0f / 64, frac);
    const float d2 = (0.5f * d) * d;
    const float s = brp_lut::kTable[2 * seg], c = brp_lut::kTable[2 * seg + 1];
    return std::fma(-d2, s, std::fma(d, c, s));
}

// log of the upper tail of chi^2 with 2k degrees of freedom:
// Q(x; 2k) = e^{-x/2} sum_{i<k} (x/2)^i / i!, summed in the log domain.
double log_upper_tail(double x, uint32_t k) {
    if (x <= 0) return 0.0;
    const double half = 0.5 * x, lh = std::log(half);
    double logs[64];
    double top = -INFINITY, lfact = 0;
    for (uint32_t i = 0; i < k; ++i) {
        if (i) lfact += std::log(double(i));
        logs[i] = double(i) * lh - lfact;
        top = std::max(top, logs[i]);
    }
    double acc = 0;
    for (uint32_t i = 0; i < k; ++i) acc += std::exp(logs[i] - top);
    return top + std::log(acc) - half;
}

double upper_tail(double x, uint32_t k) {
    const double l = log_upper_tail(x, k);
    return l < -745.0 ? 0.0 : std::exp(l);
}

// Quantile of the upper tail: the x with Q(x; 2k) = q, by bisection on the monotone
// log tail (the role gsl_cdf_chisq_Qinv plays at src/workload.cpp:122).  Every template
// of one series length asks for the same five quantiles, so the last answer per degree
// count is kept.
double upper_tail_quantile(double q, uint32_t k) {
    thread_local double memo_q[17] = {}, memo_x[17] = {};
    if (k <= 16 && memo_q[k] == q) return memo_x[k];
    const double goal = std::log(q);
    double lo = 0, hi = 1;
    while (log_upper_tail(hi, k) > goal) hi *= 2;
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        (log_upper_tail(mid, k) > goal ? lo : hi) = mid;
    }
    const double x = 0.5 * (lo + hi);
    if (k <= 16) { memo_q[k] = q; memo_x[k] = x; }
    return x;
}

uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

namespace brp {

// The fields of src/science.cpp:46-65 that do not need the resampled length: validation,
// FFT shape, tau, omega, psi, dt, 1/dt and the double-sine S0 (steps left 0).
int template_fields(uint32_t n, uint32_t unpadded, float tau, float period, float psi, float dt, brp_params* out) {
    const uint32_t radix = is_pow2(n) ? 1 : ((n % 3 == 0 && is_pow2(n / 3)) ? 3 : 0);
    if (!radix || n < 64 || n > 3u * (1u << 22) || unpadded < 32 || unpadded > n || dt <= 0 ||
        period <= 0 || tau < 0 || !std::isfinite(tau + period + psi + dt))
        return brp::fail(BRP_E_INVALID, "invalid observation/template");
    brp_params p{};
    p.n = n;
    p.unpadded = unpadded;
    p.fft_n = n;
    p.fft_radix = radix;
    p.fft_base = n / radix;
    p.logn = uint32_t(__builtin_ctz(p.fft_base));
    p.tau = tau;
    p.omega = float(kTwoPi / double(period));
    p.psi = psi;
    p.dt = dt;
    p.step_inv = 1.0f / dt;
    // The historical driver takes S0 from a double sine, the resampler a LUT sine.
    p.s0 = double_sine_s0(p);
    if (!std::isfinite(p.omega) || !std::isfinite(p.step_inv) || !std::isfinite(p.s0) ||
        !std::isfinite(p.tau * p.step_inv) || !std::isfinite(float(unpadded) * p.dt) ||
        !std::isfinite(p.omega * (float(unpadded) * p.dt) + p.psi))
        return brp::fail(BRP_E_INVALID, "template exceeds finite FP32 arithmetic domain");
    *out = p;
    return BRP_OK;
}

float double_sine_s0(const brp_params& p) {
    return float(double(p.tau) * std::sin(double(p.psi)) * double(p.step_inv));
}

// src/workload.cpp:109-112 without the resampled length (steps left 0): the search
// parameters' check, the template fields, the FMA arithmetic flags and the
// single-precision-sine S0.  The double-sine length parameters() computes first
// (src/workload.cpp:110) is the caller's to check.
int search_fields(uint32_t n, uint32_t samples, double sample_us, const brp_orbit& orbit, double f0,
                  double false_alarm, uint32_t window, brp_params* out) {
    if (!std::isfinite(f0 + false_alarm) || f0 <= 0 || false_alarm <= 0 || false_alarm >= 1 || !window)
        return brp::fail(BRP_E_INVALID, "invalid search parameters");
    brp_params p{};
    if (int rc = template_fields(n, samples, float(orbit.tau), float(orbit.period), float(orbit.psi),
                                 float(sample_us * 1e-6), &p))
        return rc;
    p.capacity = 65536;  // parameters()' result capacity (src/science.cpp:46-65)
    p.flags = 2 | 4;     // production FMA arithmetic
    p.s0 = (p.tau * std::sin(p.psi)) * p.step_inv;  // single-precision sine here
    *out = p;
    return BRP_OK;
}

// src/workload.cpp:114-124: window, F, H and the chi^2 thresholds, which depend on the
// series length and the search settings only.
int search_limits(double sample_us, double f0, double false_alarm, uint32_t window, brp_params* p) {
    const double seconds = double(p->n) * sample_us * 1e-6;
    p->window = uint32_t(window * 0.5 + 0.5);
    if (p->window >= p->n / 2 + 1) return brp::fail(BRP_E_INVALID, "search window exceeds spectrum");
    const double limit = double(p->n / 2 + 1 - p->window);
    p->fundamental = uint32_t(std::min(limit, f0 * seconds + 0.5));
    p->harmonic = uint32_t(std::min(limit, 16 * f0 * seconds + 0.5));
    const float probability = float(1.0 - std::pow(1.0 - false_alarm, 1.0 / double(p->n / 2 + 1)));
    if (!(probability > 0 && probability < 1))
        return brp::fail(BRP_E_INVALID, "false-alarm probability outside representable domain");
    for (uint32_t h = 0; h < 5; ++h)
        p->threshold[h] = float(0.5 * upper_tail_quantile(double(probability), 1u << h));
    if (p->fundamental <= p->window) return brp::fail(BRP_E_INVALID, "empty frequency search");
    return BRP_OK;
}

bool steps_jump_start_safe(const brp_params& p) {
    const double delay = std::fabs(double(p.tau) * double(p.step_inv));  // tau/dt in samples
    const double max_phase = std::fabs(double(p.omega)) * double(p.unpadded) * std::fabs(double(p.dt)) +
                             std::fabs(double(p.psi));
    return double(p.tau) * double(p.omega) < 0.5 && delay * (2.5e-5 + 1e-6 * (1.0 + max_phase)) <= 15.0;
}

}  // namespace brp

extern "C" {

float brp_lut_sin(float phase) { return lut_sin(phase); }

float brp_sample_offset(uint32_t i, const brp_params* p) {
    const float t = float(i) * p->dt;
    if (p->flags & 2) return std::fma(p->tau * fma_sin(std::fma(p->omega, t, p->psi)), p->step_inv, -p->s0);
    return p->tau * lut_sin(p->omega * t + p->psi) * p->step_inv - p->s0;
}

// The largest steps <= unpadded - 1 with float(steps) - offset(steps) < unpadded - 1 in
// the FP32 rule (src/science.cpp:41-45).  The reference scans down from unpadded - 1,
// up to ~2 tau/dt iterations.  Here the scan starts 64 samples above the FP64 solution
// s* of s - offset(s) = unpadded - 1 when that is provably safe: s - offset(s) grows with
// slope 1 - tau omega cos > 0.5, so an FP32-vs-FP64 offset difference of e samples moves
// the solution by at most 2e, and no s more than 2e above s* can satisfy the rule.  The
// offset difference is at most (tau/dt) x E with E the FP32 LUT sine's error against
// sin: the Taylor remainder |d|^3/6 <= 1.98e-5 at |d| <= pi/64, the table's 6-digit
// rounding 5e-7 x (1 + |d| + d^2/2), the reduction by 6.283185 instead of 2 pi (3.1e-7 per
// whole cycle of the phase) and the FP32 roundings of the phase (a few ulp of |phase|):
// E <= 2.5e-5 + 1e-6 (1 + |phase|).  With (tau/dt) E above 15 samples (tau/dt beyond about
// 5e5 samples) the reference's own scan from unpadded - 1 runs instead.
int brp_resampling_steps(brp_params* p) {
    if (!p || p->unpadded < 2) return brp::fail(BRP_E_INVALID, "invalid template");
    const float last = float(p->unpadded - 1);
    uint32_t steps = p->unpadded - 1;
    if (brp::steps_jump_start_safe(*p)) {
        auto offset = [&](double s) {
            return double(p->tau) * std::sin(double(p->omega) * (s * double(p->dt)) + double(p->psi)) *
                       double(p->step_inv) - double(p->s0);
        };
        // Fixed point s = (unpadded - 1) + offset(s): a contraction with factor tau omega < 1/2.
        double s = double(p->unpadded - 1);
        for (int it = 0; it < 200; ++it) {
            const double next = double(p->unpadded - 1) + offset(s);
            const bool done = std::fabs(next - s) < 1e-3;
            s = next;
            if (done) break;
        }
        const double start = std::floor(s) + 64.0;
        if (start < double(steps)) steps = start < 1.0 ? 1u : uint32_t(start);
    }
    while (steps && float(steps) - brp_sample_offset(steps, p) >= last) --steps;
    p->steps = steps;
    if (!steps) return brp::fail(BRP_E_INVALID, "template produces empty series");
    return BRP_OK;
}

int brp_parameters(uint32_t n, uint32_t unpadded, float tau, float period, float psi, float dt,
                   brp_params* out) {
    if (!out) return brp::fail(BRP_E_INVALID, "null output");
    brp_params p{};
    if (int rc = brp::template_fields(n, unpadded, tau, period, psi, dt, &p)) return rc;
    if (int rc = brp_resampling_steps(&p)) return rc;
    p.harmonic = n / 2;
    p.fundamental = n / 2;
    p.window = 8;
    const float thr[5] = {20, 30, 45, 70, 110};
    std::copy(thr, thr + 5, p.threshold);
    p.capacity = 65536;
    *out = p;
    return BRP_OK;
}

int brp_search_parameters(uint32_t n, uint32_t samples, double sample_us, const brp_orbit* orbit,
                          double f0, double false_alarm, uint32_t window, brp_params* out) {
    if (!orbit || !out) return brp::fail(BRP_E_INVALID, "null argument");
    brp_params p{};
    if (int rc = brp::search_fields(n, samples, sample_us, *orbit, f0, false_alarm, window, &p)) return rc;
    // parameters() computes the length with the double-sine S0 first (src/workload.cpp:110).
    brp_params first = p;
    first.flags = 0;
    first.s0 = brp::double_sine_s0(p);
    if (int rc = brp_resampling_steps(&first)) return rc;
    // Then with the FMA arithmetic and the single-precision sine (src/workload.cpp:111-113).
    if (int rc = brp_resampling_steps(&p)) return rc;
    if (int rc = brp::search_limits(sample_us, f0, false_alarm, window, &p)) return rc;
    *out = p;
    return BRP_OK;
}

// ---------------------------------------------------------------- toplist --
int brp_toplist_create(uint32_t rows, brp_toplist** out) {
    if (!out || !rows || rows > 100000) return brp::fail(BRP_E_INVALID, "invalid toplist size");
    auto* t = new brp_toplist;
    t->rows = rows;
    t->table.assign(5 * size_t(rows), brp_bank_row{});
    *out = t;
    return BRP_OK;
}

void brp_toplist_destroy(brp_toplist* t) { delete t; }

void brp_toplist_thresholds(const brp_toplist* t, brp_params* p) {
    for (uint32_t h = 0; h < 5; ++h)
        p->threshold[h] = std::max(p->threshold[h], float(t->table[(h + 1) * size_t(t->rows) - 1].power));
}

int brp_toplist_retain(brp_toplist* t, const brp_candidate* cand, size_t count, const brp_orbit* orbit,
                       const brp_params* p) {
    const uint32_t rows = t->rows;
    for (size_t n = 0; n < count; ++n) {
        const brp_candidate c = cand[n];
        if (c.level >= 5 || c.bin < p->window || c.bin >= p->fundamental || !std::isfinite(c.power))
            return brp::fail(BRP_E_INVALID, "invalid compact candidate");
        brp_bank_row* level = t->table.data() + size_t(c.level) * rows;
        brp_bank_row& lowest = level[rows - 1];
        if (c.power <= p->threshold[c.level] || c.power <= lowest.power) continue;
        // A bin already listed on this level is replaced only by a stronger detection.
        brp_bank_row* slot = &lowest;
        for (uint32_t i = 0; i < rows; ++i)
            if (level[i].bin == c.bin) { slot = level[i].power < c.power ? &level[i] : nullptr; break; }
        if (!slot) continue;
        // The reference rewrites the row and re-sorts the level with qsort (glibc's, a stable
        // merge sort for this size).  The level is sorted before the rewrite, so the same
        // result is one stable move of the rewritten row: out of its place, then in after
        // every row of greater power and every row of equal power that stood before it.
        const uint32_t place = uint32_t(slot - level);
        const brp_bank_row row{double(c.power), double(float(orbit->period)), double(float(orbit->tau)),
                               double(float(orbit->psi)), 0.0, 1u << c.level, c.bin};
        std::move(level + place + 1, level + rows, level + place);  // close the gap
        uint32_t at = 0;
        while (at < rows - 1 && (level[at].power > row.power || (level[at].power == row.power && at < place))) ++at;
        std::move_backward(level + at, level + rows - 1, level + rows);
        level[at] = row;
    }
    return BRP_OK;
}

size_t brp_toplist_final(const brp_toplist* t, brp_bank_row* out, size_t cap) {
    std::vector<brp_bank_row> rows = t->table;
    for (auto& r : rows) {
        if (!r.harmonics) { r.significance = -10; continue; }
        const double q = upper_tail(2 * r.power, r.harmonics);
        r.significance = q > 0 ? -std::log10(q) : 320;
        r.power /= double(float(std::sqrt(float(r.harmonics))));
    }
    auto by_sig = [](const brp_bank_row& a, const brp_bank_row& b) { return a.significance > b.significance; };
    std::stable_sort(rows.begin(), rows.end(), by_sig);
    size_t n = 0;
    while (n < t->rows && rows[0].significance > 0) {
        if (n < cap) out[n] = rows[0];
        ++n;
        const uint32_t bin = rows[0].bin;
        for (auto& r : rows) if (r.bin == bin) r.significance = -10;
        std::stable_sort(rows.begin(), rows.end(), by_sig);
    }
    return n;
}

size_t brp_toplist_rows(const brp_toplist* t, brp_bank_row* out, size_t cap) {
    const size_t n = t->table.size();
    if (out) std::copy_n(t->table.begin(), std::min(n, cap), out);
    return n;
}

int brp_write_candidates(const char* path, const brp_bank_row* rows, size_t count, double seconds) {
    std::FILE* f = std::fopen(path, "w");
    if (!f) return brp::fail(BRP_E_RUNTIME, "cannot write offline candidates");
    std::fprintf(f, "%% Experimental offline BRP4 results; never submit to production\n");
    for (size_t i = 0; i < count; ++i) {
        const brp_bank_row& c = rows[i];
        std::fprintf(f, "%6.12f %6.12f %6.12f %6.12f %g %g %d\n", c.bin / seconds, c.period, c.tau, c.psi,
                     c.power, c.significance, int(c.harmonics));
    }
    std::fprintf(f, "%%DONE%%\n");
    if (std::fclose(f)) return brp::fail(BRP_E_RUNTIME, "candidate write failed");
    return BRP_OK;
}

// -------------------------------------------------------------- synthetic --
double brp_uniform(uint64_t seed, uint64_t counter) {
    const uint64_t z = mix64(mix64(seed + 0x9E3779B97F4A7C15ull) + counter * 0xD1B54A32D192ED03ull);
    return double(z >> 11) * 0x1.0p-53;
}

int brp_synthetic_series(uint32_t count, double dt, uint64_t seed, double mean, double sigma,
                         double spin_hz, double duty, double amplitude, const brp_orbit* orbit,
                         float* out) {
    if (!out || !(dt > 0) || !(sigma >= 0)) return brp::fail(BRP_E_INVALID, "invalid synthetic series");
    const bool pulsed = amplitude != 0 && spin_hz > 0;
    double omega = 0, tau = 0, psi = 0;
    if (pulsed && orbit) {
        omega = kTwoPi / orbit->period;
        tau = orbit->tau;
        psi = orbit->psi;
    }
    const double sin_psi = std::sin(psi);
    for (uint32_t j = 0; j < count; j += 2) {
        const uint64_t pair = j / 2;
        const double u1 = brp_uniform(seed, 2 * pair), u2 = brp_uniform(seed, 2 * pair + 1);
        const double radius = std::sqrt(-2.0 * std::log(1.0 - u1));
        const double g[2] = {radius * std::cos(kTwoPi * u2), radius * std::sin(kTwoPi * u2)};
        for (uint32_t q = 0; q < 2 && j + q < count; ++q) {
            double v = mean + sigma * g[q];
            if (pulsed) {
                const double t_obs = double(j + q) * dt;
                double tp = t_obs;  // fixed point of tp = t_obs + tau (sin(omega tp + psi) - sin psi)
                for (int it = 0; it < 8; ++it) tp = t_obs + tau * (std::sin(omega * tp + psi) - sin_psi);
                const double phase = spin_hz * tp - std::floor(spin_hz * tp);
                if (phase < duty) v += amplitude * sigma;
            }
            out[j + q] = float(v);
        }
    }
    return BRP_OK;
}

int brp_synthetic_dd(uint32_t samples, double dt, uint64_t seed, double mean, double sigma, uint32_t bits,
                     double spin_hz, double duty, double amplitude, const brp_orbit* orbit, uint32_t n,
                     uint32_t rfi_count, double rfi_sigma, uint8_t* bytes, uint32_t* rfi_bins) {
    if (!bytes || (bits != 4 && bits != 8) || (bits == 4 && (samples & 1)) || samples > n || n < 64)
        return brp::fail(BRP_E_INVALID, "invalid synthetic DD block");
    std::vector<float> x(samples);
    if (int rc = brp_synthetic_series(samples, dt, seed, mean, sigma, spin_hz, duty, amplitude, orbit, x.data()))
        return rc;
    // Narrowband RFI: a cosine at bin k of the n-point spectrum whose power there is
    // rfi_sigma^2 times the mean noise power per bin (amplitude 2 rfi_sigma sigma / sqrt(n)).
    const uint64_t rseed = seed ^ 0x5bf03635d1f6a0c9ull;
    for (uint32_t r = 0; r < rfi_count; ++r) {
        const uint32_t lo = n / 64, hi = n / 2 - n / 64;
        const uint32_t k = lo + uint32_t(brp_uniform(rseed, 2ull * r) * double(hi - lo));
        const double phase = kTwoPi * brp_uniform(rseed, 2ull * r + 1);
        const double a = 2.0 * rfi_sigma * sigma / std::sqrt(double(n));
        if (rfi_bins) rfi_bins[r] = k;
        for (uint32_t j = 0; j < samples; ++j)
            x[j] = float(double(x[j]) + a * std::cos(kTwoPi * double(uint64_t(k) * j % n) / double(n) + phase));
    }
    for (uint32_t j = 0; j < samples; ++j) {
        const double q = std::nearbyint(double(x[j]));
        if (bits == 4) {
            const uint32_t v = uint32_t(std::min(15.0, std::max(0.0, q)));
            if (j & 1) bytes[j / 2] = uint8_t(bytes[j / 2] | v);
            else bytes[j / 2] = uint8_t(v << 4);  // high nibble first (src/workload.cpp:53)
        } else {
            bytes[j] = uint8_t(int8_t(std::min(127.0, std::max(-128.0, q))));
        }
    }
    return BRP_OK;
}

int brp_synthetic_bank(uint64_t seed, uint32_t count, double pmin, double pmax, double tau_max,
                       brp_orbit* out) {
    if (!out || !(pmin > 0) || !(pmax >= pmin) || !(tau_max >= 0))
        return brp::fail(BRP_E_INVALID, "invalid synthetic bank");
    const double lmin = std::log(pmin), lmax = std::log(pmax);
    for (uint32_t k = 0; k < count; ++k) {
        const uint64_t c = 3ull * k;
        out[k].period = std::exp(lmin + brp_uniform(seed, c) * (lmax - lmin));
        out[k].tau = brp_uniform(seed, c + 1) * tau_max;
        out[k].psi = brp_uniform(seed, c + 2) * kTwoPi;
    }
    return BRP_OK;
}

}  // extern "C"
