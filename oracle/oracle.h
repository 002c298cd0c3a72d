/*
 * oracle.h -- CPU restatement of the BRP4 per-template search path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may use it, and only as the checker / CPU baseline.
 *
 * Every function cites the reference file:line it restates.  The reference's
 * own CPU path (src/science.cpp + FFTW3f + the historical boinc-app-eah-brp
 * sources) cannot be built in this image (FFTW3f, GSL and the historical
 * sources are absent), so the oracle is pinned by the checks listed in
 * tests/test_oracle.py instead (closed forms, invariants, brute force,
 * numpy/scipy cross-checks).
 *
 * Arithmetic rules: compiled with -ffp-contract=off, so every FP32 expression
 * rounds exactly as written, like the reference's CMake rule
 * (CMakeLists.txt:14).  std::fma is used only where the reference calls it.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Byte-for-byte the reference Params block (include/brp/science.hpp:8-17). */
typedef struct oracle_params {
    uint32_t n, unpadded, steps, logn;
    float tau, omega, psi, dt, step_inv, s0;
    uint32_t fundamental, harmonic, window, stage;
    float threshold[5];
    uint32_t capacity;
    uint32_t fft_n, fft_base, fft_radix, flags, twiddle3, twiddle_real;
    uint32_t result_slot;
} oracle_params;

/* include/brp/science.hpp:18 */
typedef struct oracle_candidate { uint32_t bin, level; float power; uint32_t reserved; } oracle_candidate;

/* src/workload.hpp:20 BankCandidate */
typedef struct oracle_bank_row {
    double power, period, tau, psi, significance;
    uint32_t harmonics, bin;
} oracle_bank_row;

/* ---- resampling (src/science.cpp:20-45, 75-85) ---- */
float oracle_lut_sin(float phase);
float oracle_fused_sin(float phase);
float oracle_sample_offset(uint32_t i, const oracle_params* p);
int oracle_resampling_steps(oracle_params* p);
int oracle_parameters(uint32_t n, uint32_t unpadded, float tau, float period, float psi, float dt,
                      oracle_params* out);
/* resampled: n floats (steps copies + mean pad).  index: steps int32 (may be NULL). */
int oracle_resample(const float* input, size_t count, const oracle_params* p, float* resampled,
                    int32_t* index);

/* ---- spectra ---- */
/* FP64 real FFT: out = 2*(n/2+1) doubles (re, im) of sum_t x[t] e^{-2 pi i t k / n}. */
int oracle_rfft_f64(const double* x, size_t n, double* out);
/* FP32 real FFT used only for the CPU baseline (not FFTW): out = 2*(n/2+1) floats. */
int oracle_rfft_f32(const float* x, size_t n, float* out);
/* Normalised power (src/science.cpp:114-118): P[k] = |X_k|^2 / n, P[0] = 0, k < n/2.
 * Computed from the FP64 FFT of the FP32 series; power64 may be NULL. */
int oracle_power(const float* resampled, size_t n, float* power, double* power64);

/* ---- harmonic summing ---- */
/* Gather form, shaders/harmonic.comp:5-36 with shaders/hs_sum.glsl:2-10 ordering.
 * summed: 5*fundamental floats. */
int oracle_harmonic(const float* power, size_t count, const oracle_params* p, float* summed);
/* Scatter form, shaders/hs_init.comp + hs_scatter.comp + hs_collect.comp. */
int oracle_harmonic_scatter(const float* power, size_t count, const oracle_params* p, float* summed);
/* src/science.cpp:106-112; returns the number of candidates (may exceed cap). */
size_t oracle_select(const float* summed, const oracle_params* p, oracle_candidate* out, size_t cap);

/* Whole per-template pipeline (src/science.cpp:113-120) with the FP64 FFT. */
int oracle_reference(const float* input, size_t count, const oracle_params* p, float* resampled,
                     float* power, float* summed);

/* ---- bank search semantics (src/workload.cpp:108-159) ---- */
double oracle_chisq_Q(double x, double dof);     /* even dof only */
double oracle_chisq_Qinv(double q, double dof);  /* even dof only */
int oracle_search_parameters(uint32_t n, uint32_t samples, double sample_us, double period,
                             double tau, double psi, double f0, double false_alarm, uint32_t window,
                             oracle_params* out);
/* Toplist with `rows` rows per level (reference: 100). state: 5*rows entries, zeroed by caller. */
void oracle_toplist_thresholds(const oracle_bank_row* state, uint32_t rows, oracle_params* p);
int oracle_toplist_retain(oracle_bank_row* state, uint32_t rows, const oracle_candidate* c,
                          size_t count, double period, double tau, double psi, const oracle_params* p);
size_t oracle_toplist_final(const oracle_bank_row* state, uint32_t rows, oracle_bank_row* out,
                            size_t cap);

/* The bank loop's per-template work (src/workunit_main.cpp:77-99) over `ntemplates`
 * templates on `threads` threads: resample -> FP64 power (rounded to float) -> harmonic ->
 * select, each template with its own thresholds as given (the bank's base thresholds).
 * fft32 != 0 takes the FP32 FFT and power of the CPU baseline instead of the FP64 ones.
 * Candidates of template t go to out[t * cap .. t * cap + counts[t]).  A raised threshold
 * only removes candidates (a level's value is its best sum when that passes the threshold),
 * so replaying oracle_toplist_thresholds + oracle_toplist_retain over these lists in
 * template order is the reference's one-template-at-a-time loop.  Fails if a template has
 * more than cap candidates. */
int oracle_bank_candidates(const float* input, size_t count, const oracle_params* templates,
                           size_t ntemplates, uint32_t threads, uint32_t fft32, oracle_candidate* out,
                           size_t cap, uint32_t* counts);

/* ---- CPU baseline (methodology of src/reference_adapter.cpp:30-56) ---- */
/* Runs resample -> FP32 FFT -> power -> harmonic -> select for `templates` templates on
 * `threads` threads (one template per thread at a time), `reps` timed repetitions after one
 * warm-up.  Returns the median wall time of one repetition in ms (all templates). */
double oracle_benchmark_cpu(const float* input, size_t count, const oracle_params* templates,
                            size_t ntemplates, uint32_t threads, uint32_t reps);

/* ---- observation preprocessing (src/workload.cpp:29-95) ---- */
/* Decode (4-bit high nibble first, or signed 8-bit; float(double(v) / scale)),
 * zero-pad to n, FP64 real FFT, power float(re^2 + im^2), running median of `window` bins
 * (even window: mean of the two middle values), optional RFI replacement (power above
 * rfi_ratio x median scaled down to the median), whitening by float(sqrt(ln 2 / median)) on
 * [window/2, n/2 + 1 - window/2) with every other bin zero, FP64 inverse FFT over the full
 * Hermitian spectrum, divided by sqrt(n).  out: samples floats; decoded (may be NULL):
 * samples floats; rfi_bins (may be NULL): replaced bins. */
int oracle_preprocess(const uint8_t* bytes, size_t samples, uint32_t bits, double scale, uint32_t n,
                      uint32_t window, float rfi_ratio, float* out, float* decoded, uint32_t* rfi_bins);
/* oracle_preprocess with the zap list of src/workload.cpp:76-86 (zap_ranges (lower, upper)
 * Hz pairs, bins uint32(f n sample_us 1e-6 + 0.5), inclusive, replaced after whitening by
 * Gaussian noise of sigma sqrt(2)/2 sqrt(padding) from oracle_uniform; band edges stay
 * zero).  The reference draws GSL taus2 ziggurat noise, which is absent: the noise
 * values are parity unpinned, the bins, sigma and edges are checked. */
int oracle_preprocess_zap(const uint8_t* bytes, size_t samples, uint32_t bits, double scale, uint32_t n,
                          uint32_t window, float rfi_ratio, const double* zap, size_t zap_ranges, double sample_us,
                          double padding, float* out, float* decoded, uint32_t* rfi_bins);
/* Counter-based uniform in [0, 1): SplitMix64 finaliser of seed and counter. */
double oracle_uniform(uint64_t seed, uint64_t counter);

/* The 65-entry (sin, cos) table used by the LUT sine, as 130 interleaved floats. */
void oracle_lut_table(float* out);

const char* oracle_last_error(void);

#ifdef __cplusplus
}
#endif
