// engine.h -- interfaces between the handle (engine.hip) and the kernel files.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "brp_hip.h"

namespace brp_engine {

// ----------------------------------------------------------------- plain --
struct PlainBuffers {
    uint32_t n = 0, fft_n = 0, packed = 0;
    const float* x = nullptr;        // observation, unpadded samples
    const float2* lut = nullptr;     // 65 (sin, cos) segments
    float* raw = nullptr;            // resampled samples, i < steps
    int32_t* index = nullptr;        // resample source index (snapshot only, else null)
    double* partial = nullptr;       // per-block FP64 sums
    float* mean = nullptr;           // pad value
    float2* a = nullptr;             // FFT ping
    float2* scratch = nullptr;       // FFT pong
    float2* X = nullptr;             // n/2 + 1 spectrum bins
    const float2* tw_fft = nullptr;  // W_fft_n^t, t < fft_n
    const float2* tw_real = nullptr; // W_n^k, k <= n/2
};

void plain_clear_slots(hipStream_t st, uint32_t* results, uint32_t slot_words, uint32_t nslots);
void plain_resample(hipStream_t st, const PlainBuffers& b, const brp_params& p, uint32_t* slot);
float2* plain_fft(hipStream_t st, const PlainBuffers& b, bool radix8, bool radix4);
void plain_spectrum(hipStream_t st, const PlainBuffers& b, const float2* z, bool fused_power, float* power);
void plain_power(hipStream_t st, const PlainBuffers& b, float* power);
void plain_harmonic(hipStream_t st, const float* power, float* summed, const brp_params& p, bool scatter,
                    uint32_t* slot);

// ----------------------------------------------------------------- fused --
// The half-length complex transform of M = N/2 points is split M = R x C.  Pass 1: for
// each row r, resample+pack z[r + R c] (c < C), FFT over c, multiply by W_M^{r k2}, store
// Y[r][k2].  Pass 2: for each column k2 (with its mirror column C - k2), FFT over r,
// real split, pad correction, power P[k1 C + k2].  Harmonic: dyadic sums + candidates.
struct FusedShape {
    uint32_t n = 0, m = 0, c = 0, r = 0;  // N, M = N/2, pass-1 length C, pass-2 length R
    uint32_t s1 = 0, s2 = 0;              // rows per pass-1 block, lower columns per pass-2 block
    uint32_t t1 = 0, t2 = 0;              // threads per pass-1 / pass-2 block
    uint32_t blocks1 = 0, blocks2 = 0;    // blocks per template
    uint32_t mirror = 0;                  // exchange rows with each column beside its mirror
    uint32_t n_shift = 0, n_odd = 0;      // N = n_odd 2^n_shift, n_odd 1 or 3
};

struct FusedArgs {
    const brp_params* params = nullptr;  // per template slot (device)
    const float* x = nullptr;
    const float2* lut = nullptr;
    const float2* tw_c = nullptr;    // pass-1 per-stage twiddle table (fft_stage_table(C, ...))
    const float2* tw_r = nullptr;    // pass-2 per-stage twiddle table (fft_stage_table(R, ...))
    const float2* twm_lo = nullptr;  // W_M^t, t < 2048
    const float2* twm_hi = nullptr;  // W_M^{2048 q}
    const float2* tw_rq = nullptr;   // W_{16 R}^t, t < 16 R
    const float2* twn_lo = nullptr;  // W_N^t, t < 2048
    const float2* twn_hi = nullptr;  // W_N^{2048 q}
    float2* Y = nullptr;             // per group slot: M complex (stride M of the group's series)
    float* P = nullptr;              // per group slot: M floats
    double* partial = nullptr;       // per group slot: blocks1 doubles
    float* mean = nullptr;           // per group slot
    uint32_t* results = nullptr;     // per batch slot
    uint32_t slot_words = 0;
    uint32_t first = 0;              // batch slot of group slot 0
    uint32_t count = 0;              // templates in this group
    uint32_t fp64_phase = 0;
    uint32_t fma = 0;                // flag 2 arithmetic (uniform over a handle)
    uint32_t nonneg = 0;             // every template of the group has omega, psi, dt >= 0
    uint32_t precision = 0;          // split-FP16 MFMA GEMMs: bit 0 pass-2, bit 1 pass-1 radix-16 stages
    uint32_t power_bins = 0;         // pass 2 writes power bins [0, power_bins) only (0: all M)
    // snapshot (per batch slot, or null)
    float* snap_resampled = nullptr;  // N floats
    int32_t* snap_index = nullptr;    // N int32
    float2* snap_X = nullptr;         // M + 1 complex
    float* snap_summed = nullptr;     // 5 F floats
    uint32_t snap_f = 0;              // summed stride (largest F)
    uint32_t snap_n = 0;              // resampled / index stride (largest N)
    uint32_t snap_m1 = 0;             // spectrum stride (largest M + 1)
};

// The shape of series length n with pass-1 length c, pass-1 rows per block s1, pass-2
// column pairs per block s2 and threads per block t1, t2 (each 0 = the first match).
bool fused_supported(uint32_t n, uint32_t c, uint32_t s1, uint32_t s2, FusedShape* shape, uint32_t t1 = 0,
                     uint32_t t2 = 0);
// Whether the shape has the radix-16 stages the precision variant runs as MFMA GEMMs: for
// bit 0 a pass-2 plan starting with radix 16, for bit 1 a pass-1 shape compiled with them.
bool fused_mfma_capable(const FusedShape& s, uint32_t precision);
// The per-stage twiddle table of a fused FFT pass (length len, `elements` per thread): for
// each stage i >= 1 of the radix plan (radix[i], ns[i] = product of the earlier radices) a
// block at offset[i] whose entry (q - 1) ns[i] + k, q in [1, radix[i]), k < ns[i], holds
// W_len^{k q len / (ns[i] radix[i])}; count entries in all.
struct StageTable {
    uint32_t len = 0, count = 0, stages = 0;
    uint32_t offset[4] = {0, 0, 0, 0}, ns[4] = {0, 0, 0, 0}, radix[4] = {0, 0, 0, 0};
    // The exponent of entry e (e < count).
    __host__ __device__ uint64_t exponent(uint32_t e) const {
        uint32_t i = 0;
        while (i + 1 < stages && e >= offset[i + 1]) ++i;
        const uint32_t within = e - offset[i], q = within / ns[i] + 1, k = within % ns[i];
        return uint64_t(k) * q * (len / (ns[i] * radix[i]));
    }
};
StageTable fft_stage_table(uint32_t len, uint32_t elements);
uint32_t fused_pass1_elements(const FusedShape& s);  // C S1 / T1
uint32_t fused_pass2_elements(const FusedShape& s);  // 2 R S2 / T2
// Every compiled shape of series length n, the default first.
std::vector<FusedShape> fused_shapes(uint32_t n);
// One launch each covers all `count` templates of the group.
void fused_pass1(hipStream_t st, const FusedShape& s, const FusedArgs& a);
void fused_pass2(hipStream_t st, const FusedShape& s, const FusedArgs& a);
// The harmonic kernel over bins [0, reach) of every template of the group (reach: the bins
// the fundamentals below F need, at most min(16 F, H + 8)).
void fused_harmonic(hipStream_t st, const FusedShape& s, const FusedArgs& a, uint32_t reach);
// The power bins [0, result) the harmonic kernel reads for that reach (whole tiles): the
// FusedArgs::power_bins a group without snapshots passes to pass 2.
uint32_t harmonic_power_bins(uint32_t reach);
// resampling_steps of every template of params[0 .. count) whose steps is 0, on the
// device (the host's jump-started scan, for templates where it is safe); error bit 1 when
// one produces an empty series.
void search_steps(hipStream_t st, brp_params* params, uint32_t count, const float2* lut, uint32_t* error);
// bytes / 16 uint4 copies from src to dst (a multiple of 16 bytes).
void copy16(hipStream_t st, const void* src, void* dst, size_t bytes);
// --------------------------------------------------------- device toplist --
// One toplist row on the device (the host brp_bank_row's fields as the retain writes them:
// float values, significance 0).  Tables are [toplist][level][row].
struct TopRow {
    float power, period, tau, psi;
    uint32_t bin, harmonics;
};
constexpr uint32_t kDeviceToplistMaxRows = 2048;
// p.threshold[h] = max(p.threshold[h], lowest[table * 5 + h]) for `count` slots
// (toplist.hip); lowest holds each level's last-row power as a retain recorded it.
void toplist_thresholds(hipStream_t st, brp_params* params, uint32_t count, const uint32_t* slot_table,
                        const float* lowest);
// Toplist::retain of `count` result slots in slot order into the device tables, then the
// last-row power of each level into lowest[table * 5 + h]; error bits 1 (slot invalid or
// overflowed), 2 (invalid candidate), 4 (too many candidates of a level).
void toplist_retain(hipStream_t st, const uint32_t* results, uint32_t slot_words, const brp_params* params,
                    uint32_t count, const uint32_t* slot_table, const float4* slot_orbit, uint32_t tables,
                    TopRow* rows_dev, uint32_t rows, float* lowest, uint32_t* error);

// ------------------------------------------------------------ preprocess --
// Decode + whitening + RFI replacement + zap list of one observation (preprocess.hip);
// throws std::runtime_error on an invalid median or zap range.  zap: zap_ranges pairs
// (lower, upper) in Hz; sample_us and padding size the zap bins and noise.  decoded (may
// be null) receives the decoded samples, rfi_bins (may be null) the number of replaced
// bins.
void preprocess_observation(const uint8_t* bytes, size_t samples, uint32_t bits, double scale, uint32_t n,
                            uint32_t window, float rfi_ratio, const double* zap, size_t zap_ranges, double sample_us,
                            double padding, float* out, float* decoded, uint32_t* rfi_bins);

// Copies {count, flags} and the first `keep` candidates of each of `nslots` slots into a
// dense staging area of (keep + 1) x 16 bytes per slot.
void pack_results(hipStream_t st, const uint32_t* results, uint32_t slot_words, uint32_t nslots, uint32_t keep,
                  uint32_t* staging);

}  // namespace brp_engine
