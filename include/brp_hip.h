/*
 * brp_hip.h -- C ABI of the MI355X (gfx950) BRP4 per-template search backend.
 *
 * This is the drop-in boundary.  It replaces the reference's Vulkan backend
 * class brp::GPU (include/brp/vulkan.hpp:6-29, src/vulkan.cpp:278-406) and the
 * host science helpers that feed it (include/brp/science.hpp:8-38,
 * include/brp/workload.hpp:6-24).  Plain pointers and sizes only; no C++ or torch
 * types cross it.  include/brp/hip.hpp wraps it in the reference's C++ class shape.
 *
 * Errors: every function returning int returns 0 on success and a nonzero
 * BRP_E_* class otherwise; brp_last_error() then holds a message for the calling
 * thread (the reference threw std::runtime_error with the same messages).
 *
 * Threading: one handle per host thread; calls on one handle must not overlap.
 */
#ifndef BRP_HIP_H
#define BRP_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- types ---- */

/* Per-template parameter block, byte-compatible with the reference Params
 * (include/brp/science.hpp:8-17, 108 bytes).  Flag bits (src/vulkan.cpp:111,
 * 279-282): 1 packed real FFT, 2 production FMA arithmetic in the resampler,
 * 4 V3D FMA emulation (exact FMA is native here, so 4 changes nothing),
 * 8 scatter harmonic search, 16 radix-8 stages, 32 radix-8 paired store,
 * 64 fused power. */
typedef struct brp_params {
    uint32_t n, unpadded, steps, logn;
    float tau, omega, psi, dt, step_inv, s0;
    uint32_t fundamental, harmonic, window, stage;
    float threshold[5];
    uint32_t capacity;
    uint32_t fft_n, fft_base, fft_radix, flags, twiddle3, twiddle_real;
    uint32_t result_slot;
} brp_params;

/* include/brp/science.hpp:18 */
typedef struct brp_candidate { uint32_t bin, level; float power; uint32_t reserved; } brp_candidate;

/* Orbit of the bank (include/brp/workload.hpp:13). */
typedef struct brp_orbit { double period, tau, psi; } brp_orbit;

/* Toplist row (include/brp/workload.hpp:15). */
typedef struct brp_bank_row {
    double power, period, tau, psi, significance;
    uint32_t harmonics, bin;
} brp_bank_row;

/* Engines. */
enum {
    BRP_ENGINE_PLAIN = 0, /* reference-shaped: one template per launch sequence,
                             one kernel per stage (resample, mean, pack, Stockham
                             passes, real split, power, harmonic) */
    BRP_ENGINE_FUSED = 1  /* two-pass four-step FFT with resampling fused into pass 1
                             and real split + power fused into pass 2; all templates
                             of a batch in each launch; hipGraph replay */
};

/* Constructor options.  The first eleven fields are the reference GPU constructor's
 * arguments (include/brp/vulkan.hpp:14) with the same meaning and validation
 * (src/vulkan.cpp:116-122, 279-283). */
typedef struct brp_options {
    uint32_t workgroup;    /* 16..256; plain engine block size */
    uint32_t tile;         /* 0 or 64..4096 power of two (reference LDS DIT tile); accepted,
                              validated, and mapped onto Stockham stages */
    uint32_t batch;        /* templates per run (result slots); 1..4096 */
    uint32_t software;     /* accepted for API compatibility; no effect */
    uint32_t stockham, radix4, packed, scatter, radix8, fused_power, radix8_store;
    uint32_t engine;       /* BRP_ENGINE_* */
    uint32_t graph;        /* fused engine: replay the batch as a captured hipGraph */
    uint32_t snapshot;     /* keep resampled / FFT / summed arrays for snapshot() */
    uint32_t fp64_phase;   /* evaluate the orbital offset in FP64 (variant; indices may
                              differ from the FP32 LUT rule, which is the parity mode) */
    uint32_t split_c;      /* fused engine: pass-1 FFT length C (0 = default table) */
    uint32_t precision;    /* 0 = FP32 FFT; split-FP16 MFMA radix-16 stages: 1 = pass 2's,
                              2 = pass 1's (the transform's first stages), 3 = both */
    uint32_t group;        /* fused engine: templates per launch (0 = 8) */
    uint32_t split_s1;     /* fused engine: pass-1 rows per block (0 = default for split_c) */
    uint32_t split_s2;     /* fused engine: pass-2 column pairs per block (0 = default) */
    uint32_t lanes;        /* fused engine: groups in flight (1 or 2 streams; 0 = 1), each lane
                              with its own exchange and power buffers */
    uint32_t untimed;      /* fused engine: no per-kernel events (stage times read as 0) */
    uint32_t timing_stride; /* fused engine: per-kernel events on every n-th launch group only
                               (0 or 1 = every group); stage times and brp_hip_kernel_ms then
                               cover the timed groups */
    uint32_t split_t1;     /* fused engine: threads per pass-1 block (0 = the first shape that fits) */
    uint32_t split_t2;     /* fused engine: threads per pass-2 block (0 = the first shape that fits) */
} brp_options;

/* src/vulkan.hpp:6 Timings; template_stage_ms is returned separately. */
typedef struct brp_timings {
    double wall_ms, cpu_ms, submit_ms;
    double stage_ms[4]; /* resample_prepare, fft, power, harmonic_candidates */
} brp_timings;

typedef struct brp_hip brp_hip;

enum { BRP_OK = 0, BRP_E_INVALID = 1, BRP_E_RUNTIME = 2, BRP_E_OVERFLOW = 3, BRP_E_DEVICE = 4 };

const char* brp_last_error(void);

/* ------------------------------------------------------- host science ------- */

/* src/science.cpp:20-28 */
float brp_lut_sin(float phase);
/* src/science.cpp:36-40 */
float brp_sample_offset(uint32_t i, const brp_params* p);
/* src/science.cpp:41-45 */
int brp_resampling_steps(brp_params* p);
/* src/science.cpp:46-65 (bench semantics) */
int brp_parameters(uint32_t n, uint32_t unpadded, float tau, float period, float psi, float dt,
                   brp_params* out);
/* src/workload.cpp:108-125 (bank semantics); the chi^2 quantile uses the closed form of
 * the even-dof upper tail in place of GSL's gsl_cdf_chisq_Qinv. */
int brp_search_parameters(uint32_t n, uint32_t samples, double sample_us, const brp_orbit* orbit,
                          double f0, double false_alarm, uint32_t window, brp_params* out);

/* Toplist (src/workload.cpp:126-159), rows per level configurable (reference: 100). */
typedef struct brp_toplist brp_toplist;
int brp_toplist_create(uint32_t rows_per_level, brp_toplist** out);
void brp_toplist_destroy(brp_toplist* t);
void brp_toplist_thresholds(const brp_toplist* t, brp_params* p);
int brp_toplist_retain(brp_toplist* t, const brp_candidate* c, size_t count, const brp_orbit* orbit,
                       const brp_params* p);
/* Final significance-ranked rows; returns the row count (written up to cap). */
size_t brp_toplist_final(const brp_toplist* t, brp_bank_row* out, size_t cap);
/* Raw 5 x rows table (level-major, power-descending within a level). */
size_t brp_toplist_rows(const brp_toplist* t, brp_bank_row* out, size_t cap);
/* src/workload.cpp:160-165 results file ("%DONE%" terminated). */
int brp_write_candidates(const char* path, const brp_bank_row* rows, size_t count, double seconds);

/* ----------------------------------------------- synthetic workloads -------- */
/* Seeded N(0, sigma^2) noise (counter-based RNG + Box-Muller in double) plus a
 * rectangular pulse train (frequency, duty, amplitude in units of sigma) whose pulse
 * times are Doppler-shifted by a circular orbit: pulsar-frame time t_p of sample j
 * solves t_p - tau sin(omega t_p + psi) + tau sin psi = j dt.  amplitude 0 gives pure
 * noise.  Matches the resampler's model, so the matching template makes the train
 * periodic. */
int brp_synthetic_series(uint32_t count, double dt, uint64_t seed, double mean, double sigma,
                         double spin_hz, double duty, double amplitude, const brp_orbit* orbit,
                         float* out);
/* A DD sample block (the bytes after the header, src/workload.cpp:45-53): the series of
 * brp_synthetic_series (mean, sigma, pulse train) plus rfi_count narrowband cosines at
 * seeded bins k in [n/64, n/2 - n/64) of the n-point spectrum with power rfi_sigma^2 times
 * the mean noise power there, rounded to the nearest integer and clamped: bits 4 ->
 * [0, 15], two per byte, high nibble first; bits 8 -> signed [-128, 127].  rfi_bins (may
 * be NULL) receives the rfi_count bins. */
int brp_synthetic_dd(uint32_t samples, double dt, uint64_t seed, double mean, double sigma, uint32_t bits,
                     double spin_hz, double duty, double amplitude, const brp_orbit* orbit, uint32_t n,
                     uint32_t rfi_count, double rfi_sigma, uint8_t* bytes, uint32_t* rfi_bins);
/* Orbits drawn from a seed: P log-uniform in [pmin, pmax], tau uniform in [0, tau_max],
 * psi uniform in [0, 2 pi). */
int brp_synthetic_bank(uint64_t seed, uint32_t count, double pmin, double pmax, double tau_max,
                       brp_orbit* out);
/* Counter-based uniform in [0, 1) (SplitMix64 of seed and counter). */
double brp_uniform(uint64_t seed, uint64_t counter);

/* ---------------------------------------------------- preprocessing -------- */
/* load_observation + preprocess (src/workload.cpp:29-95) on the device, for the sample
 * block of a DD file (after its header): decode (bits = 4: packed, high nibble first;
 * bits = 8: signed bytes; each value float(double(v) / scale)), zero-pad to n, real FFT,
 * power float(re^2 + im^2), running median of `window` bins (even, 2 <= window <= 3072,
 * window <= n / 2; even-window median = mean of the two middle values), whitening by
 * sqrt(ln 2 / median) with the bins outside [window/2, n/2 + 1 - window/2) zeroed, and
 * inverse FFT scaled by 1/sqrt(n).  RFI (new behaviour; the reference zaps a listed range
 * with GSL noise): a bin whose power exceeds rfi_ratio x its running median (rfi_ratio 0 =
 * off) is scaled to the median first.  out: `samples` floats (the reference's
 * Observation::preprocessed); decoded (may be NULL): `samples` floats; rfi_bins (may be
 * NULL): bins replaced.  Errors: invalid arguments, or a non-positive median
 * ("invalid whitening median", as src/workload.cpp:68). */
int brp_hip_preprocess(const uint8_t* bytes, size_t samples, uint32_t bits, double scale, uint32_t n,
                       uint32_t window, float rfi_ratio, float* out, float* decoded, uint32_t* rfi_bins);
/* brp_hip_preprocess with the reference's zap list (src/workload.cpp:76-86): zap holds
 * zap_ranges (lower, upper) pairs in Hz; after whitening, the bins uint32(f T + 0.5),
 * T = n * sample_us * 1e-6, of each range (inclusive) are replaced by Gaussian noise of
 * sigma sqrt(2)/2 * sqrt(padding) per component, except the band edges, which stay zero.
 * The noise is this library's (Box-Muller from brp_uniform, seeded by the bit pattern of
 * the first decoded sample, counters 2k and 2k + 1 for bin k), not GSL's taus2 ziggurat.
 * Errors as the reference: "invalid zap range" (non-finite, lower < 0 or upper < lower)
 * and "zap exceeds Nyquist" (last bin > n/2). */
int brp_hip_preprocess_zap(const uint8_t* bytes, size_t samples, uint32_t bits, double scale, uint32_t n,
                           uint32_t window, float rfi_ratio, const double* zap, size_t zap_ranges, double sample_us,
                           double padding, float* out, float* decoded, uint32_t* rfi_bins);

/* ------------------------------------------------------------ device -------- */

/* GPU(const Params&, span<const float> input, ...) -- include/brp/vulkan.hpp:14.
 * Copies `count` == p->unpadded input samples to the device. */
int brp_hip_create(const brp_params* p, const float* input, size_t count, const brp_options* opt,
                   brp_hip** out);
int brp_hip_destroy(brp_hip* h);
/* set_templates -- src/vulkan.cpp:289-295.  count <= options.batch.  Templates must keep
 * n, unpadded, fundamental, harmonic and capacity, and the arithmetic flags (2|4). */
int brp_hip_set_templates(brp_hip* h, const brp_params* p, size_t count);
/* run -- src/vulkan.cpp:296-318.  Blocks until the batch is done.  template_stage_ms
 * (may be NULL) receives 4 doubles per template. */
int brp_hip_run(brp_hip* h, brp_timings* t, double* template_stage_ms);
/* candidates(slot) -- src/vulkan.cpp:329-337: sorted by (level, bin).  Returns
 * BRP_E_OVERFLOW on candidate overflow or an out-of-range resample index.  *count gets
 * the number of candidates even when cap is too small (then nothing beyond cap is
 * written and BRP_E_INVALID is returned). */
int brp_hip_candidates(brp_hip* h, uint32_t slot, brp_candidate* out, size_t cap, size_t* count);
/* snapshot() -- src/vulkan.cpp:319-328, for the given slot (needs options.snapshot for the
 * fused engine).  Any pointer may be NULL.  resampled: n floats; index: steps int32
 * (resample source index); fft: 2*(n/2+1) floats; power: n/2 floats; summed:
 * 5*fundamental floats. */
int brp_hip_snapshot(brp_hip* h, uint32_t slot, float* resampled, int32_t* index, float* fft,
                     float* power, float* summed);
/* device_info / allocated_bytes / profile_counters -- src/vulkan.cpp:338-406.  Each writes at
 * most cap - 1 characters and a terminating NUL, and returns the full length.
 * profile_counters: JSON of per-kernel hardware counters, collected on every call by replaying
 * the handle's last batch once per counter set under rocprofiler-sdk's dispatch counting service
 * (with a plain copy of known size that calibrates FETCH_SIZE / WRITE_SIZE to bytes).  Needs
 * BRP_HIP_COUNTERS=1 in the environment when the library loads (before the ROCm runtime starts
 * in the process); otherwise, or when another tool such as rocprofv3 holds the profiler, the
 * JSON is {"available": false, "reason": ...}.  h may be NULL for that availability answer. */
size_t brp_hip_device_info(brp_hip* h, char* buf, size_t cap);
uint64_t brp_hip_allocated_bytes(const brp_hip* h);
size_t brp_hip_profile_counters(brp_hip* h, char* buf, size_t cap);
/* Per-stage device times (ms) accumulated over every run since the previous call of this
 * function (which resets them), measured with HIP events on the launch stream:
 * resample/pass 1, fft/pass 2, power, harmonic.  *launches (may be NULL) receives the
 * number of launch sequences timed: groups of templates (fused engine; with
 * options.timing_stride n, every n-th group) or templates (plain engine); *templates (may be
 * NULL) the templates those launches covered. */
int brp_hip_kernel_ms(brp_hip* h, double out[4], uint64_t* launches, uint64_t* templates);
/* Autotuner (no reference counterpart; the reference picks its radix fusion and tile by
 * hand-run sweeps).  For the series length of templates[0], times the fused engine on the
 * caller's `count` templates (one run of `count` per measurement, `reps` measurements
 * after a warm-up, median kept) for every compiled four-step split (pass-1 length C and
 * rows per block S1, pass-2 length R and column pairs per block S2: the LDS tiles, block
 * sizes and radix sequences), every templates-per-launch value in {2, 4, 8, 16} not above
 * count, and 1 or 2 lanes (only base->lanes when that is nonzero).
 * *best receives `base` with split_c, split_s1, split_s2, split_t1, split_t2, group and lanes
 * (1, 2) set to the fastest; report (may be NULL)
 * receives one JSON object per candidate, newline separated. */
int brp_hip_autotune(const brp_params* templates, size_t count, const float* input, size_t input_count,
                     const brp_options* base, uint32_t reps, brp_options* best, char* report, size_t cap);
/* Measurement aid with no reference counterpart: copies `bytes` between two device
 * buffers `reps` times with a plain streaming kernel (brp_copy) and returns the mean
 * GB/s; used to calibrate hardware byte counters against a transfer of known size. */
int brp_hip_copy_bandwidth(size_t bytes, int reps, double* gbps);

/* Uneven batches (the batched extension of set_templates, src/vulkan.cpp:289-295, which
 * lifts the reference's fixed-dimension rule): registers one more observation of
 * p->unpadded samples, padded to p->n (2^k or 3 2^k), on a fused-engine handle.  A
 * template then names its series by (n, unpadded); templates of every registered series
 * can share one batch, one launch schedule and one graph, each with its own F and H.
 * The capacity and arithmetic flags stay those of the handle.  Option overrides of the
 * four-step split apply where they fit the length, else its default shape is used. */
int brp_hip_add_series(brp_hip* h, const brp_params* p, const float* input, size_t count);
/* Bank search over prepared templates (brp_search_parameters or brp_parameters results,
 * any registered series), the bank loop of src/workunit_main.cpp:77-136 with one toplist
 * per series: tops[i] collects the templates of series i (registration order, the
 * constructor's series first), ntops >= the number of series named.  Thresholds are
 * frozen per batch and retain() replayed in template order within each series (the bank
 * runs series by series, so batches are whole launch groups and graphs are replayed), so
 * each toplist equals the one-template-at-a-time result for its series.  cand_count (may
 * be NULL) receives the candidate count of each template. */
int brp_hip_bank_search_templates(brp_hip* h, const brp_params* templates, const brp_orbit* orbits, size_t count,
                                  brp_toplist* const* tops, size_t ntops, uint32_t* cand_count, double* gpu_ms);

/* brp_hip_bank_search_templates with the toplists on the device (the next step after
 * compaction: Toplist::thresholds / retain, src/workload.cpp:126-141, run in the launch
 * stream after each batch).  The toplists are uploaded at the start and written back at
 * the end; they equal brp_hip_bank_search_templates' row for row.  Fused engine without
 * snapshots; every toplist needs the same number of rows, at most 2048 per level; at most
 * 4096 candidates of one level per template.  Candidates stay on the device; cand_count
 * (may be NULL) still receives each template's count. */
int brp_hip_bank_search_device(brp_hip* h, const brp_params* templates, const brp_orbit* orbits, size_t count,
                               brp_toplist* const* tops, size_t ntops, uint32_t* cand_count, double* gpu_ms);
/* Toplist::retain (src/workload.cpp:129-141) of `count` templates on the device, for
 * candidates produced elsewhere: cand_count[i] candidates of template i, concatenated in
 * `candidates` (any order within a template), each template retained with its own
 * parameters and orbit, in order.  Equals brp_toplist_thresholds + brp_toplist_retain per
 * template.  Same limits as brp_hip_bank_search_device. */
int brp_hip_toplist_retain_device(const brp_params* templates, const brp_orbit* orbits, size_t count,
                                  const brp_candidate* candidates, const uint32_t* cand_count, brp_toplist* top);

/* Bank search: the brp_workunit bank loop (src/workunit_main.cpp:77-136) for a bank of
 * `count` orbits against the handle's series, in batches of options.batch templates,
 * with thresholds frozen at each batch start and Toplist::retain replayed in template
 * order, so the toplist equals the one-template-at-a-time result.  Template parameters
 * come from brp_search_parameters(n, samples, sample_us, ...).  templates_out (may be
 * NULL) receives the count parameter blocks as run; cand_count (may be NULL) the
 * candidate count of each template. */
int brp_hip_bank_search(brp_hip* h, const brp_orbit* orbits, size_t count, uint32_t samples,
                        double sample_us, double f0, double false_alarm, uint32_t window,
                        brp_toplist* top, brp_params* templates_out, uint32_t* cand_count,
                        double* gpu_ms);

/* --------------------------------------------------- no allocation -------- */
/* No counterpart in the reference.  For callers that cannot allocate device memory: code
 * under HIP graph capture, serving loops and harnesses that forbid hipMalloc.  Runs the fused
 * engine's pipeline (resample + pass 1, pass 2 + power, harmonic sums + candidates) for
 * `count` templates (1 .. BRP_NOALLOC_MAX_TEMPLATES) of one series, with every intermediate
 * buffer (parameters, LUT, twiddle tables, exchange, power, partial sums) in a file-scope
 * device array of BRP_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.  The
 * templates must share n, unpadded, capacity and the arithmetic flags; N x count must fit the
 * array (N = 2^22: 2 templates; N = 2^20: 8), else BRP_E_INVALID with the sizes.
 *   input:   device pointer to the series (input_count >= unpadded + 2 readable floats; the
 *            two after the series are read and never used);
 *   results: device pointer to count result slots of 16 + 16 * capacity bytes each, filled as
 *            the handle's slots are ({count, flags, 0, 0}, then candidates in no order; flag
 *            1 = candidate overflow, 2 = a resample index outside the series);
 *   stream:  a hipStream_t (NULL: the default stream).  The call only enqueues work.
 * Nothing is allocated and nothing is copied from host memory, so the call can be captured.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU race. */
#define BRP_NOALLOC_SCRATCH_BYTES (64u << 20)
#define BRP_NOALLOC_MAX_TEMPLATES 8
int brp_hip_search_noalloc(const brp_params* templates, size_t count, const float* input, size_t input_count,
                           uint32_t* results, void* stream);

#ifdef __cplusplus
}
#endif
#endif
