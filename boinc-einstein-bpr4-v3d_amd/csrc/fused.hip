// fused.hip -- the fused engine for gfx950.
//
// Per template the real series of N samples is packed into M = N/2 complex points and
// transformed as a two-pass four-step FFT with M = R x C:
//
//   pass 1  (block = S1 consecutive rows r, all c):  resample + pack z[r + R c] straight
//           from the observation (the bit-exact index rule of src/science.cpp:75-85),
//           FFT over c in registers/LDS, multiply by W_M^{r k2}, store Y[r][k2] (rows are
//           contiguous, so every store is a full line).  The pad is left at zero here and
//           each block writes an FP64 partial sum of its samples.
//   pass 2  (block = S2 columns k2 plus their mirror columns C - k2, all r):  FFT over r,
//           giving Z[k1 C + k2] and Z[M - k] in the same block, then the real split
//           X[k] = 1/2 [(Z_k + conj Z_{M-k}) - i W_N^k (Z_k - conj Z_{M-k})], the mean pad
//           added by linearity, X += mu U with U[k] = sum_{t >= steps} W_N^{tk}
//           = (W_N^{k steps} - 1) / (1 - W_N^k), and the power |X|^2 / N.
//   harmonic (block = one wave, 16 bins per lane):  dyadic sums of the lane's bins in
//           registers from one run of consecutive powers per harmonic, then the maximum over
//           each fundamental's rounding interval (8 sums from the left lane by a wave shift),
//           threshold test and candidate append (the gather result of
//           shaders/harmonic.comp:14-36 without recomputing sums per level).
//
// One launch of each kernel covers every template of a group, blocks mapped so that
// consecutive work items stay on one XCD (8 XCDs, block b runs on XCD b mod 8).
#include "device.h"
#include "engine.h"

#pragma clang fp contract(off)

using namespace brp_dev;
using brp_engine::FusedArgs;
using brp_engine::FusedShape;

namespace {


struct Radices { int count; int r[5]; };

constexpr Radices radices_for(int L) {
    switch (L) {
        case 64: return {2, {16, 4}};
        case 128: return {2, {16, 8}};
        case 256: return {2, {16, 16}};
        case 512: return {3, {8, 8, 8}};
        case 1024: return {3, {16, 16, 4}};
        case 2048: return {3, {16, 16, 8}};
        case 4096: return {3, {16, 16, 16}};
        case 384: return {3, {3, 16, 8}};
        case 768: return {3, {3, 16, 16}};
        case 1536: return {4, {3, 8, 8, 8}};
        case 3072: return {4, {3, 16, 16, 4}};
        default: return {0, {0}};
    }
}

// The radix plan of an L-point transform whose threads hold E elements: the table plan
// when every radix divides E, else the factor 3 first and then radix 16, 8, 4 or 2 as
// E allows (for example 8 8 8 8 for L = 4096, E = 8).
constexpr Radices plan_for(int L, int E) {
    const Radices base = radices_for(L);
    bool ok = base.count > 0;
    int prod = 1;
    for (int i = 0; i < base.count; ++i) {
        if (E % base.r[i]) ok = false;
        prod *= base.r[i];
    }
    if (ok && prod == L) return base;
    Radices p{0, {0, 0, 0, 0, 0}};
    int left = L;
    if (left % 3 == 0 && E % 3 == 0) { p.r[p.count++] = 3; left /= 3; }
    while (left > 1 && p.count < 5) {
        const int r = (left % 16 == 0 && E % 16 == 0) ? 16
                      : (left % 8 == 0 && E % 8 == 0) ? 8
                      : (left % 4 == 0 && E % 4 == 0) ? 4 : 2;
        p.r[p.count++] = r;
        left /= r;
    }
    return left == 1 ? p : Radices{0, {0, 0, 0, 0, 0}};
}

template <int L, int E, int ST>
constexpr int ns_before() {
    int p = 1;
    for (int i = 0; i < ST; ++i) p *= plan_for(L, E).r[i];
    return p;
}

constexpr bool plan_ok(int L, int e) {
    const Radices pl = plan_for(L, e);
    int prod = 1;
    for (int i = 0; i < pl.count; ++i) {
        if (e % pl.r[i]) return false;
        prod *= pl.r[i];
    }
    return pl.count > 0 && prod == L;
}

// Offset of stage ST's block in the per-stage twiddle table of an L-point transform at E
// elements per thread (layout: brp_engine::fft_stage_table).
template <int L, int E, int ST>
constexpr int stage_table_offset() {
    const Radices pl = plan_for(L, E);
    int off = 0, ns = pl.r[0];
    for (int i = 1; i < ST; ++i) {
        off += (pl.r[i] - 1) * ns;
        ns *= pl.r[i];
    }
    return off;
}

__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t grid) {
    if (grid % 8u) return bid;
    return (bid % 8u) * (grid / 8u) + bid / 8u;
}

// LDS slot of element l of sequence s: sequences interleaved, element-major, with one
// padding row per 16 elements.  The first Stockham stage stores element 16 j + q from
// lane j, a stride of 16 S elements; the padding turns that into 17 S, so the lanes of
// a wave fall into different LDS banks.
template <int S>
__device__ __forceinline__ int lds_at(int l, int s) { return (l + (l >> 4)) * S + s; }
template <int L, int S>
constexpr int lds_size() { return (L + L / 16) * S; }

// Stockham stage ST of an L-point FFT over S sequences held by T threads, E = L S / T
// elements each.  On entry v[u R + q] holds input q of butterfly u = (s, j), b = t + T u,
// s = b mod S, j = b / S, i.e. element j + q L/R.  On exit of the last stage v[u R + q]
// holds output element j + q L/R of sequence s.  tw is the transform's per-stage table:
// input q of a butterfly with k = j mod NS takes W_L^{k q L / (NS R)} from entry
// (q - 1) NS + k of the stage's block, so for each q a wave reads a run of consecutive
// entries (one or two cache lines) rather than entries k q apart.
// MF: radix-16 stages as split-FP16 MFMA GEMMs (the precision variant; wave-collective,
// so every lane of the block runs the same stages).
template <int L, int S, int T, int ST, bool MF = false>
__device__ __forceinline__ void fft_stages(cf* v, float2* lds, const float2* __restrict__ tw, int t,
                                           const Dft16Mfma* mf = nullptr) {
    constexpr int E = L * S / T;
    constexpr Radices pl = plan_for(L, E);
    constexpr int R = pl.r[ST];
    constexpr int NS = ns_before<L, E, ST>();
    constexpr int U = E / R;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int b = t + T * u;
        const int j = b / S;
        const int k = j % NS;
        if constexpr (NS > 1) {
#pragma unroll
            for (int q = 1; q < R; ++q)
                v[u * R + q] = mul(v[u * R + q], load(tw + (stage_table_offset<L, E, ST>() + (q - 1) * NS + k)));
        }
        if constexpr (MF && R == 16) dft16_mfma(v + u * R, *mf);
        else dft<R>(v + u * R);
    }
    if constexpr (ST + 1 < pl.count) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int b = t + T * u;
            const int s = b % S, j = b / S, k = j % NS;
            const int out = (j / NS) * NS * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) store(lds + lds_at<S>(out + q * NS, s), v[u * R + q]);
        }
        __syncthreads();
        constexpr int R2 = pl.r[ST + 1];
        constexpr int U2 = E / R2;
#pragma unroll
        for (int u = 0; u < U2; ++u) {
            const int b = t + T * u;
            const int s = b % S, j = b / S;
#pragma unroll
            for (int q = 0; q < R2; ++q) v[u * R2 + q] = load(lds + lds_at<S>(j + q * (L / R2), s));
        }
        fft_stages<L, S, T, ST + 1, MF>(v, lds, tw, t, mf);
    }
}

__device__ __forceinline__ cf twiddle2(const float2* __restrict__ lo, const float2* __restrict__ hi, uint32_t e) {
    return mul(load(hi + (e >> 11)), load(lo + (e & 2047u)));
}

// Position of column k2 in a row of the exchange Y.  Plain layout: k2.  Mirror layout
// (FusedShape::mirror): every column beside its mirror C - k2 (0 and C/2 first: 0, C/2, 1,
// C-1, 2, C-2, ...), so the columns a pass-2 block pairs up form one contiguous run of each
// row.  That halves the segments pass 2 reads when it pairs few columns per block, at the
// price of half-filled store instructions in pass 1; the shape table turns it on for
// shapes of at most 4 column pairs per pass-2 block, where it measured faster.
__device__ __forceinline__ uint32_t exchange_pos(uint32_t k2, uint32_t C) {
    const uint32_t h = C / 2;
    return k2 == 0 ? 0u : (k2 == h ? 1u : (k2 < h ? 2u * k2 : 2u * (C - k2) + 1u));
}

// ------------------------------------------------------------------ pass 1 --
// Resample + pack the E elements of thread t: element e (butterfly u = e / R0, input
// q = e % R0) is the packed pair z[m] = (y[2m], y[2m+1]) with m = r0 + s + R c,
// c = j + q C/R0.  The two
// samples of a pair come from source indices j0 <= j1 <= j0 + 2 (the orbit moves the
// index by far less than one sample per sample), so one 12-byte load at
// b = clamp(j0, 0, U - 1) serves both (the observation buffer holds zero guard samples
// after U - 1); any other pair takes a second load.  Samples outside the observation
// are masked to zero afterwards and flagged.
// NN: phase >= 0 for every sample; FM: flag 2 arithmetic; F64: FP64 orbital phase; ALL:
// every sample of the block lies below steps (all blocks but the last few of a template).
template <int C, int S, int T, bool SNAP, bool F64, bool NN, bool FM, bool ALL>
__device__ __forceinline__ void resample_pairs(cf* v, double& acc, bool& bad, const FusedArgs& a,
                                               const FusedShape& sh, const brp_params& p,
                                               const float4* __restrict__ lut4, uint32_t r0, uint32_t slot,
                                               int t) {
    constexpr int E = C * S / T;
    constexpr int R0 = plan_for(C, E).r[0];
    const uint32_t R = sh.r;
    const uint32_t steps = p.steps, U = p.unpadded;
    const Orbit32 o{p.tau, p.omega, p.psi, p.dt, p.step_inv, p.s0, p.flags};
    int tm = t;  // the thread's index as element_m reads it
    auto element_m = [&](int e) -> uint32_t {
        const int b = tm + T * (e / R0);
        const int s = b % S, j = b / S;
        return r0 + s + __umul24(R, uint32_t(j + (e % R0) * (C / R0)));  // R, c < 2^24
    };
    auto index_of = [&](uint32_t i) -> int {
        if constexpr (F64) {
            const double tau_dt = double(p.tau) / double(p.dt);
            const double s0_dt = double(p.tau) * sin_f64(double(p.psi)) / double(p.dt);
            return source_index_fp64(i, tau_dt, double(p.omega), double(p.psi), double(p.dt), s0_dt);
        } else {
            return source_index4<NN, FM>(i, o, lut4);
        }
    };
    // The unmasked form's index: equal to index_of whenever it lies in [0, U) (U < 2^24), the
    // only case in which that form keeps its result.
    auto index_fast = [&](uint32_t i) -> int {
        if constexpr (F64) return index_of(i);
        else return source_index4<NN, FM, true>(i, o, lut4);
    };
    // Validation runs: the pairs of elements [e0, e0 + count) and their source indices
    // (computed again) for the samples below steps.
    auto write_snapshot = [&](int e0, int count) {
        if constexpr (SNAP) {
#pragma unroll
            for (int c = 0; c < count; ++c) {
                const uint32_t i = 2 * element_m(e0 + c);
                if (ALL || i < steps) {
                    a.snap_resampled[size_t(slot) * a.snap_n + i] = v[e0 + c].x;
                    a.snap_index[size_t(slot) * a.snap_n + i] = index_of(i);
                }
                if (ALL || i + 1 < steps) {
                    a.snap_resampled[size_t(slot) * a.snap_n + i + 1] = v[e0 + c].y;
                    a.snap_index[size_t(slot) * a.snap_n + i + 1] = index_of(i + 1);
                }
            }
        }
    };
    // Up to 16 elements at once (all of them at 16 or fewer elements per thread; 24-element
    // plans in two halves), so two selector bits per element fit one word.
    constexpr int CH = E <= 16 ? E : E / 2;
    static_assert(E % CH == 0 && CH <= 16, "gather chunks");
    struct Trio { float v[3]; };
    // Interior blocks (every sample below steps) first take the form without per-sample
    // masks: per element only the selector bits, and per lane the smallest first and the
    // largest second index.  When every lane of the wave stayed inside the observation, no
    // sample needs masking; otherwise the wave starts again in the masked form below (rare:
    // only orbits that reach past the observation's ends).  Both forms add the same values
    // in the same order, so the sums are identical.
    if constexpr (ALL && CH == E) {
        Trio w[E];
        uint32_t sel = 0;
        int lo = 0x7fffffff, hi = -0x7fffffff;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t m = element_m(e);
            const int j0 = index_fast(2 * m), j1 = index_fast(2 * m + 1);
            // An address inside the observation for any j0 (a negative j0 reads at U - 1 and
            // sends the wave to the masked form anyway).
            const uint32_t b = min(uint32_t(j0), U - 1);
            __builtin_memcpy(&w[e], reinterpret_cast<const char*>(a.x) + b * 4u, sizeof(Trio));
            sel |= min(uint32_t(j1) - b, 3u) << (2 * e);
            lo = min(lo, j0);
            hi = max(hi, j1);
            asm volatile("" : "+v"(sel), "+v"(lo), "+v"(hi));
            __builtin_amdgcn_sched_barrier(0);
        }
        if (__all(lo >= 0 && hi < int(U))) {
            double sum = 0.0;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t code = (sel >> (2 * e)) & 3u;
                const uint32_t u0 = __builtin_bit_cast(uint32_t, w[e].v[0]);
                const uint32_t u1 = __builtin_bit_cast(uint32_t, w[e].v[1]);
                const uint32_t u2 = __builtin_bit_cast(uint32_t, w[e].v[2]);
                const uint32_t b0 = 0u - (code & 1u), b1 = 0u - ((code >> 1) & 1u);
                const uint32_t low = (u1 & b0) | (u0 & ~b0);
                const uint32_t pick = ((u2 & b1) | (low & ~b1)) & (0u - uint32_t(code != 3u));
                const float y0 = w[e].v[0], y1 = __builtin_bit_cast(float, pick);
                sum += double(y0);
                sum += double(y1);
                v[e] = cf{y0, y1};
            }
            if (sel & (sel >> 1) & 0x55555555u) {  // selector 3: the second sample's own load
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    if (((sel >> (2 * e)) & 3u) == 3u) {
                        uint32_t mm = element_m(e);
                        asm volatile("" : "+v"(mm));
                        const int j1 = index_of(2 * mm + 1);
                        const bool inside = uint32_t(j1) < U;
                        bad |= !inside;
                        const float y1 = inside ? a.x[j1] : 0.0f;
                        sum += double(y1);
                        v[e].y = y1;
                    }
                }
            }
            acc += sum;
            write_snapshot(0, E);
            return;
        }
        // The masked form from an opaque thread index: reusing the unmasked form's index
        // arithmetic would keep all of it live through that form's gather.
        asm volatile("" : "+v"(tm));
    }
#pragma unroll
    for (int e0 = 0; e0 < E; e0 += CH) {
        // Phase 1: per element, both source indices, then the 12-byte load at
        // b = clamp(j0, 0, U - 1) at once, so every load of the chunk is in flight while the
        // later indices are computed.  Kept per element besides the loaded trio: which word
        // of it holds the second sample (0, 1, 2; 3 = neither, a second load below) and
        // whether each sample lies in the series and inside the observation (else it is 0).
        Trio w[CH];
        uint32_t sel = 0, keep = 0, outside = 0;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const uint32_t m = element_m(e0 + c);
            const int j0 = index_of(2 * m), j1 = index_of(2 * m + 1);
            const int b = min(max(j0, 0), int(U) - 1);
            // 32-bit byte offset from the observation's base (U < 2^30 samples)
            __builtin_memcpy(&w[c], reinterpret_cast<const char*>(a.x) + uint32_t(b) * 4u, sizeof(Trio));
            const uint32_t d1 = uint32_t(j1 - b);
            sel |= min(d1, 3u) << (2 * c);
            const bool s0 = ALL || 2 * m < steps, s1 = ALL || 2 * m + 1 < steps;
            const bool in0 = uint32_t(j0) < U, in1 = uint32_t(j1) < U;
            outside |= uint32_t((s0 && !in0) || (s1 && !in1));
            keep |= (uint32_t(s0 && in0) | (uint32_t(s1 && in1) << 1)) << (2 * c);
            // The element's bits are settled here (sinking their arithmetic into phase 2
            // would keep both indices of every element live), and one element's index
            // arithmetic runs at a time (interleaving them would keep every intermediate
            // live beside the loads in flight).
            asm volatile("" : "+v"(sel), "+v"(keep), "+v"(outside));
            __builtin_amdgcn_sched_barrier(0);
        }
        bad |= outside != 0u;
        // Phase 2: the pairs from the trios, masked, and the FP64 sum of the series samples
        // in element order.
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            // Bitwise selects (v_bfi / v_and): a select between the trio's words by the
            // selector would become an indexed array in scratch memory.
            const uint32_t code = (sel >> (2 * c)) & 3u;
            const uint32_t k = keep >> (2 * c);
            const uint32_t u0 = __builtin_bit_cast(uint32_t, w[c].v[0]);
            const uint32_t u1 = __builtin_bit_cast(uint32_t, w[c].v[1]);
            const uint32_t u2 = __builtin_bit_cast(uint32_t, w[c].v[2]);
            const uint32_t b0 = 0u - (code & 1u), b1 = 0u - ((code >> 1) & 1u);
            const uint32_t low = (u1 & b0) | (u0 & ~b0);  // selector 0 or 1
            const uint32_t pick = (u2 & b1) | (low & ~b1);
            const uint32_t keep0 = 0u - (k & 1u), keep1 = 0u - (((k >> 1) & 1u) & uint32_t(code != 3u));
            const float y0 = __builtin_bit_cast(float, u0 & keep0);
            const float y1 = __builtin_bit_cast(float, pick & keep1);
            acc += double(y0);
            acc += double(y1);
            v[e0 + c] = cf{y0, y1};
        }
        // Rare: a second index not within two samples of the first (selector 3) takes its
        // own load.  The index is computed again from an opaque m: reusing the first
        // computation would keep every element's second index live through the gather.
        if (sel & (sel >> 1) & 0x55555555u) {
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                if (((sel >> (2 * c)) & 3u) == 3u && ((keep >> (2 * c)) & 2u)) {
                    uint32_t mm = element_m(e0 + c);
                    asm volatile("" : "+v"(mm));
                    const float y1 = a.x[min(max(index_of(2 * mm + 1), 0), int(U) - 1)];
                    acc += double(y1);
                    v[e0 + c].y = y1;
                }
            }
        }
        write_snapshot(e0, CH);
    }
}

// SNAP: also write the resampled samples and their source indices (validation runs).
// F64: orbital offset from the FP64 sine (a reported variant; the FP32 LUT rule is the
// parity mode).
// MF: the row FFT's radix-16 stages (the first stages of the transform) as split-FP16 MFMA
// GEMMs (the precision variant; compiled for the shapes pass1_mfma_instance names).
// At least 4 waves per SIMD (128 VGPRs): blocks of 512 threads then run two per CU; with
// eight elements per thread, 6 waves per SIMD (80 VGPRs), three 512-thread blocks per CU.
template <int C, int S, int T, bool SNAP, bool F64, bool NN, bool FM, bool MF>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(C * S / T <= 8 && !F64 ? 6 : 4))) void k_pass1(
    FusedArgs a, FusedShape sh) {
    constexpr int E = C * S / T;
    static_assert(plan_ok(C, E), "pass-1 radix plan does not fit the block");
    constexpr Radices pl = plan_for(C, E);
    constexpr int RL = pl.r[pl.count - 1];
    __shared__ float2 lds[lds_size<C, S>()];
    __shared__ float4 lut4[65];
    __shared__ double red[T / 64];
    const int t = threadIdx.x;
    const uint32_t item = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t g = item / sh.blocks1, rb = item % sh.blocks1;
    const uint32_t slot = a.first + g;
    const brp_params& p = a.params[slot];
    const uint32_t M = sh.m;
    const uint32_t r0 = rb * S;
    if (t < 65) {
        const float2 sc = a.lut[t];
        lut4[t] = make_float4(sc.x, sc.y, 0.5f * sc.x, 0.0f);
    }
    __syncthreads();

    double acc = 0.0;
    bool bad = false;
    cf v[E];
    // The block's last sample: m = r0 + S - 1 + R (C - 1), i = 2 m + 1.
    if (2u * (r0 + S - 1 + sh.r * (C - 1)) + 1u < p.steps)
        resample_pairs<C, S, T, SNAP, F64, NN, FM, true>(v, acc, bad, a, sh, p, lut4, r0, slot, t);
    else
        resample_pairs<C, S, T, SNAP, F64, NN, FM, false>(v, acc, bad, a, sh, p, lut4, r0, slot, t);
    // Keep the FFT's table reads out of the gather: hoisted there, their registers would
    // sit beside the gather's loads in flight.
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (MF) {
        const Dft16Mfma mf = dft16_mfma_setup();
        fft_stages<C, S, T, 0, true>(v, lds, a.tw_c, t, &mf);
    } else {
        fft_stages<C, S, T, 0>(v, lds, a.tw_c, t);
    }

    // Y[r][k2] = v * W_M^{r k2} for k2 = j + q C/RL: W_M^{r j} times
    // W_M^{r q C/RL} = W_{R RL}^{r q} (r q < R RL, read from the W_{16 R} table), all of a
    // thread's table reads issued together.  Column k2 goes to position exchange_pos(k2, C)
    // of its row (in the mirror layout 2 k2 below C/2, 2 (C - k2) + 1 above, C/2 at 1).
    // The block's S rows are assembled in LDS first and then written as 16-byte pieces of
    // consecutive positions, so every store instruction of a wave fills whole lines (the
    // mirror positions taken straight from registers would half-fill twice as many).
    const bool mirror = sh.mirror != 0;
    float2* stage = lds;  // S rows of C positions: S C <= lds_size<C, S>()
    __syncthreads();      // the last FFT stage's LDS reads are done
#pragma unroll
    for (int u = 0; u < E / RL; ++u) {
        const int b = t + T * u;
        const int s = b % S, j = b / S;
        const uint32_t r = r0 + s;
        cf pw[RL];
#pragma unroll
        for (int q = 1; q < RL; ++q) pw[q] = load(a.tw_rq + __umul24(r, uint32_t(q * (16 / RL))));
        const cf base = twiddle2(a.twm_lo, a.twm_hi, __umul24(r, uint32_t(j)));
        float2* row = stage + s * C;
#pragma unroll
        for (int q = 0; q < RL; ++q) {
            const cf w = q == 0 ? base : mul(base, pw[q]);
            const int k2 = j + q * (C / RL);
            const int pos = !mirror ? k2
                            : (q < RL / 2 ? 2 * k2 : ((q == RL / 2 && j == 0) ? 1 : 2 * (C - k2) + 1));
            store(row + pos, mul(v[u * RL + q], w));
        }
    }
    __syncthreads();
    {
        static_assert((S * C) % (2 * T) == 0, "16-byte row pieces per thread");
        float4* Y4 = reinterpret_cast<float4*>(a.Y + size_t(g) * M + size_t(r0) * C);  // rows r0 .. r0 + S - 1
        const float4* st4 = reinterpret_cast<const float4*>(stage);
#pragma unroll
        for (int k = 0; k < S * C / (2 * T); ++k) Y4[t + T * k] = st4[t + T * k];
    }

    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((t & 63) == 0) red[t >> 6] = acc;
    if (bad) atomicOr(a.results + size_t(slot) * a.slot_words + 1, 2u);
    __syncthreads();
    if (t == 0) {
        double total = 0.0;
        for (int w = 0; w < T / 64; ++w) total += red[w];
        a.partial[size_t(g) * sh.blocks1 + rb] = total;
    }
}

// (a b) mod N for a, b < N, N = odd 2^shift with odd 1 or 3: the low `shift` bits of a b stay
// and the rest reduces mod odd (a b < N^2, so (a b) >> shift < 3 N fits in 32 bits; hi / 3 =
// umulhi(hi, 0xAAAAAAAB) / 2 exactly for every 32-bit hi).
__device__ __forceinline__ uint32_t mulmod_n(uint32_t a, uint32_t b, const FusedShape& sh) {
    const uint64_t x = uint64_t(a) * b;
    const uint32_t lo = uint32_t(x) & ((1u << sh.n_shift) - 1u);
    uint32_t hi = uint32_t(x >> sh.n_shift);
    hi = sh.n_odd == 3u ? hi - 3u * (__umulhi(hi, 0xAAAAAAABu) >> 1) : 0u;
    return (hi << sh.n_shift) | lo;
}

// ------------------------------------------------------------------ pass 2 --
// SNAP: also write the spectrum X (validation runs).
// MF: the split-FP16 MFMA variant of the radix-16 stages (fewer waves per SIMD: the
// matrix operands and accumulators take registers).
template <int R, int S, int T, bool SNAP, bool MF>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(MF ? 2 : (2 * R * S / T <= 8 ? 8 : 4)))) void
k_pass2(FusedArgs a, FusedShape sh) {
    constexpr int SS = 2 * S;
    constexpr int E = R * SS / T;
    static_assert(plan_ok(R, E), "pass-2 radix plan does not fit the block");
    constexpr Radices pl = plan_for(R, E);
    constexpr int R0 = pl.r[0], RL = pl.r[pl.count - 1];
    __shared__ float2 lds[lds_size<R, SS>()];
    __shared__ double red[T / 64 + 1];
    const int t = threadIdx.x;
    const uint32_t item = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t g = item / sh.blocks2, w = item % sh.blocks2;
    const uint32_t slot = a.first + g;
    const brp_params& p = a.params[slot];
    const uint32_t C = sh.c, M = sh.m, N = sh.n;
    const uint32_t base = w * S;
    // Sequence s < S is the lower column base + s, sequence S + s its mirror C - base - s.
    // Columns 0 and C/2 are their own mirrors: block 0 keeps column 0 in sequence 0 and
    // takes column C/2 in sequence S, so C/(2S) blocks cover every column once.
    const bool first_block = w == 0;
    auto column = [&](int s) -> uint32_t {
        const uint32_t lower = base + uint32_t(s % S);
        if (s < S) return lower;
        return (first_block && s == S) ? C / 2 : C - lower;
    };

    // This block's share of the pass-1 partial sums, read before the columns so that the
    // mean reduction below waits only for it.
    const double* partial = a.partial + size_t(g) * sh.blocks1;
    double part = uint32_t(t) < sh.blocks1 ? partial[t] : 0.0;

    // Column reads as 32-bit offsets from the slot's base (M complex < 2^32 bytes).
    const char* Y = reinterpret_cast<const char*>(a.Y + size_t(g) * M);
    cf v[E];
#pragma unroll
    for (int u = 0; u < E / R0; ++u) {
        const int b = t + T * u;
        const int s = b % SS, j = b / SS;
        const uint32_t col = column(s);
        const uint32_t pos = sh.mirror ? exchange_pos(col, C) : col;
        const uint32_t at = (uint32_t(j) * C + pos) * uint32_t(sizeof(float2));
#pragma unroll
        for (int q = 0; q < R0; ++q) {
            const uint32_t byte = at + uint32_t(q * (R / R0)) * C * uint32_t(sizeof(float2));
            v[u * R0 + q] = load(reinterpret_cast<const float2*>(Y + byte));
        }
    }

    // Pad mean from the pass-1 partial sums, in a fixed order (bit-reproducible), while
    // the column loads are in flight.
    for (uint32_t b = t + T; b < sh.blocks1; b += T) part += partial[b];
    for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off, 64);
    if ((t & 63) == 0) red[t >> 6] = part;
    __syncthreads();
    if (t == 0) {
        double total = 0.0;
        for (int wv = 0; wv < T / 64; ++wv) total += red[wv];
        red[T / 64] = total;
    }
    __syncthreads();
    const float mu = float(red[T / 64] / double(p.steps));
    if (w == 0 && t == 0) a.mean[slot] = mu;

    if constexpr (MF) {
        const Dft16Mfma mf = dft16_mfma_setup();
        fft_stages<R, SS, T, 0, true>(v, lds, a.tw_r, t, &mf);
    } else {
        fft_stages<R, SS, T, 0>(v, lds, a.tw_r, t);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < E / RL; ++u) {
        const int b = t + T * u;
        const int s = b % SS, j = b / SS;
#pragma unroll
        for (int q = 0; q < RL; ++q) store(lds + lds_at<SS>(j + q * (R / RL), s), v[u * RL + q]);
    }
    __syncthreads();

    float* P = a.P + size_t(g) * M;
    // Power bins past the harmonic kernel's reach are not written (snapshots keep them all).
    const uint32_t lim = (SNAP || a.power_bins == 0) ? uint32_t(M) : a.power_bins;
    float2* snapX = SNAP ? a.snap_X + size_t(slot) * a.snap_m1 : nullptr;
    const float inv_n = 1.0f / float(N);
    const uint32_t steps = p.steps;
    const float pad_len = float(N - steps);
    const float parity = (steps & 1u) ? -1.0f : 1.0f;  // W_N^{M steps} = (-1)^steps
    // Pad-correction twiddle W_N^{k steps}: a thread's pairs share the column (T is a
    // multiple of S) and step k by T/S rows, so e(k) = k steps mod N advances by the
    // constant D = (T/S) C steps mod N.  W_N^{e} starts from one table lookup and is then
    // multiplied by W_N^D per pair (the correction term tolerates the few-ulp drift).
    static_assert(T % S == 0, "pair mapping");
    const uint32_t steps_n = steps < N ? steps : steps - N;  // steps <= N
    const uint32_t k_first = uint32_t(t / S) * C + base + uint32_t(t % S);
    cf we = twiddle2(a.twn_lo, a.twn_hi, mulmod_n(k_first, steps_n, sh));
    const cf wstep = twiddle2(a.twn_lo, a.twn_hi, mulmod_n(uint32_t(T / S) * C, steps_n, sh));
    // W_N^k of the thread's pairs: pair u lies (T/S) u rows after pair 0, so
    // W_N^{k_u} = W_N^{k_0} W_N^{(T/S) C u}, whose second factor is the same for every thread
    // (scalar table reads).
    constexpr int PAIRS = R * S / T;
    cf Wk[PAIRS];
    Wk[0] = twiddle2(a.twn_lo, a.twn_hi, k_first);
#pragma unroll
    for (int u = 1; u < PAIRS; ++u)
        Wk[u] = mul(Wk[0], twiddle2(a.twn_lo, a.twn_hi, uint32_t(u) * uint32_t(T / S) * C));
    // X[k] and X[M - k] from Z[k] and Z[M - k] (zk, zm), W = W_N^k, we = W_N^{k steps}:
    // X[k] = (A - i W B) / 2 and X[M - k] = conj(A + i W B) / 2 with A = Z_k + conj Z_{M-k} and
    // B = Z_k - conj Z_{M-k} (as W_N^{M-k} = -conj W_N^k), so one product W B serves both.
    auto real_split = [&](const cf zk, const cf zm, const cf W, const cf we_k, uint32_t k, cf& xk, cf& xm) {
        const cf A = add(zk, conj(zm)), B = sub(zk, conj(zm));
        const cf WB = mul(B, W);
        xk = cf{(A.x + WB.y) * 0.5f, (A.y - WB.x) * 0.5f};
        xm = cf{(A.x - WB.y) * 0.5f, (-A.y - WB.x) * 0.5f};
        if (k == 0) {
            xk.x += mu * pad_len;
            xm.x += mu * (parity - 1.0f) * 0.5f;  // U[M] = ((-1)^steps - 1) / 2
        } else {
            // cot(pi k / N) and tan(pi k / N) = cot(pi (M - k) / N) from W_N^k = (cos, -sin),
            // each in the half-angle form without cancellation for its half of (0, pi).
            const bool front = W.x >= 0.0f;
            const float num = front ? 1.0f + W.x : -W.y;
            const float den = front ? -W.y : 1.0f - W.x;
            const float ck = num * __builtin_amdgcn_rcpf(den);
            const float cm = den * __builtin_amdgcn_rcpf(num);
            const cf uk = mul(cf{we_k.x - 1.0f, we_k.y}, cf{0.5f, -0.5f * ck});
            const cf um = mul(cf{parity * we_k.x - 1.0f, -parity * we_k.y}, cf{0.5f, -0.5f * cm});
            xk = cf{__builtin_fmaf(mu, uk.x, xk.x), __builtin_fmaf(mu, uk.y, xk.y)};
            xm = cf{__builtin_fmaf(mu, um.x, xm.x), __builtin_fmaf(mu, um.y, xm.y)};
        }
    };
#pragma unroll
    for (int u = 0; u < PAIRS; ++u) {
        const int pi = t + T * u;
        const int s = pi % S;
        const uint32_t k1 = uint32_t(pi / S);
        if (u) we = mul(we, wstep);
        const uint32_t col = base + uint32_t(s);
        const bool self = col == 0;
        // Column 0 pairs k1 with R - k1 within its own sequence.
        const uint32_t k1m = self ? (R - k1) % R : R - 1 - k1;
        const int ms = self ? 0 : s + S;
        const uint32_t k = k1 * C + col;
        const cf zk = load(lds + lds_at<SS>(int(k1), s));
        const cf zm = load(lds + lds_at<SS>(int(k1m), ms));
        cf xk, xm;
        real_split(zk, zm, Wk[u], we, k, xk, xm);
        if (k < lim) P[k] = k == 0 ? 0.0f : (xk.x * xk.x + xk.y * xk.y) * inv_n;
        if (!self && M - k < lim) P[M - k] = (xm.x * xm.x + xm.y * xm.y) * inv_n;
        if constexpr (SNAP) {
            store(snapX + k, xk);
            if (!self || k == 0) store(snapX + (M - k), xm);
        }
    }
    // Column C/2 (sequence S of block 0) pairs k1 with R - 1 - k1 within itself; each
    // thread takes rows t, t + T, ... and writes X[k] only.
    if (first_block) {
        for (uint32_t k1 = uint32_t(t); k1 < R; k1 += T) {
            const uint32_t k = k1 * C + C / 2;
            const cf zk = load(lds + lds_at<SS>(int(k1), S));
            const cf zm = load(lds + lds_at<SS>(int(R - 1 - k1), S));
            const cf W = twiddle2(a.twn_lo, a.twn_hi, k);
            const cf we_k = twiddle2(a.twn_lo, a.twn_hi, mulmod_n(k, steps_n, sh));
            cf xk, xm;
            real_split(zk, zm, W, we_k, k, xk, xm);
            if (k < lim) P[k] = (xk.x * xk.x + xk.y * xk.y) * inv_n;
            if constexpr (SNAP) store(snapX + k, xk);
        }
    }
}

// ---------------------------------------------------------------- harmonic --
// Each lane owns 16 consecutive bins i = B + d (d < 16, B a multiple of 16).  The dyadic sum
// of bin i reads P[(i k + 8) >> 4] for k = 1 .. 16, and (B + d) k + 8 >> 4 = B k / 16 +
// ((d k + 8) >> 4), so for each k the lane reads one run of (15 k + 8 >> 4) + 1 consecutive
// powers from B k / 16 (144 words for its 16 bins instead of 256 single-word gathers; the
// runs of neighbouring lanes are adjacent, so a wave reads one stretch per k), and the bin
// offsets within a run are compile-time constants.
// A block is one wave: lane 0 covers the 16 bins before the tile and produces nothing (its
// last 8 sums of each level feed lane 1), lanes 1 .. 63 own the tile's 1008 bins.  The
// intervals of level h, [j 2^h - 2^(h-1), j 2^h + 2^(h-1)), start at most 8 bins before the
// owning lane's first bin: each lane takes its left neighbour's last 8 sums of the level by
// a wave shift and finds every maximum in registers.
constexpr int kBinsPerLane = 16;
constexpr int kTileBins = 63 * kBinsPerLane;  // bins owned per block

constexpr int run_width(int k) { return ((15 * k + 8) >> 4) + 1; }

template <int K>
struct Run { float v[run_width(K)]; };

template <int K>
__device__ __forceinline__ Run<K> load_run(const float* __restrict__ P, int B) {
    Run<K> r;
    __builtin_memcpy(&r, P + (B / 16) * K, sizeof(r));  // B k / 16, exact: B is a multiple of 16
    return r;
}

// Runs of k >= 5 go through LDS.  Loaded straight from P, a wave instruction of lane-owned
// runs touches lanes k words apart (about 2 k cache lines per instruction, one tag lookup
// each), and the L1 tag rate bounds the kernel.  Instead the wave loads the stretch of P its
// 64 runs cover, [s k, s k + 63 k + width), as contiguous 16-byte pieces (lane m takes words
// 4 m + 256 j), writes it to its LDS buffer and reads each lane's run from there.  A lane's
// run starts lb k words in (lb: the lane's 16-bin block from the wave's first); for k = 8, 12
// and 16 one padding word follows every k words, so that the 32 lanes of an LDS cycle read
// with an odd stride (k + 1) and hit distinct banks; the other k are odd (no conflict) or
// 2 mod 4 (two-way).  A wave reads and writes only its own buffer, and a wave executes its LDS
// instructions in order, so no barrier separates the phases.
constexpr bool staged(int k) { return k >= 5; }
constexpr bool padded(int k) { return k == 8 || k == 12 || k == 16; }
constexpr int stretch_len(int k) { return 63 * k + run_width(k); }
constexpr int pieces(int k) { return (stretch_len(k) + 255) / 256; }
constexpr int stage_words(int k) { return padded(k) ? 256 * pieces(k) + 256 * pieces(k) / k : 256 * pieces(k); }
constexpr int kStageWords = 1088;  // the largest stage_words(k), k = 5 .. 16
constexpr bool stage_fits() {
    for (int k = 5; k <= 16; ++k)
        if (stage_words(k) > kStageWords) return false;
    return true;
}
static_assert(stage_fits(), "stage buffer");

template <int K>
struct Pieces { float4 c[pieces(K)]; };

// The wave's stretch of harmonic K from P + s K.  Lanes whose piece lies past the stretch
// load the stretch's last 4 words instead (no branch, no extra cache line); their pieces
// land past the stretch in LDS, where no run reads.
template <int K>
__device__ __forceinline__ Pieces<K> fetch_stretch(const float* __restrict__ P, int s, int m) {
    Pieces<K> pc;
    const float* src = P + s * K;
#pragma unroll
    for (int j = 0; j < pieces(K); ++j)
        __builtin_memcpy(&pc.c[j], src + min(4 * m + 256 * j, stretch_len(K) - 4), 16);
    return pc;
}

// LDS word of stretch word e.
template <int K>
__device__ __forceinline__ int stage_at(int e) { return padded(K) ? e + e / K : e; }

template <int K>
__device__ __forceinline__ Run<K> stage_run(float* __restrict__ buf, const Pieces<K>& pc, int m, int lb) {
    asm volatile("" ::: "memory");  // the previous stretch's reads stay before these writes
#pragma unroll
    for (int j = 0; j < pieces(K); ++j) {
        const int e = 4 * m + 256 * j;
        if constexpr (padded(K)) {
            // The 4 words of a piece share one padding offset (e is a multiple of 4 and
            // e mod k + 3 < k for k = 8, 12, 16).
            float* d = buf + stage_at<K>(e);
            d[0] = pc.c[j].x;
            d[1] = pc.c[j].y;
            d[2] = pc.c[j].z;
            d[3] = pc.c[j].w;
        } else {
            *reinterpret_cast<float4*>(buf + e) = pc.c[j];
        }
    }
    asm volatile("" ::: "memory");
    Run<K> r;
    const float* src = buf + (padded(K) ? lb * (K + 1) : lb * K);
#pragma unroll
    for (int d = 0; d < run_width(K); ++d) r.v[d] = src[padded(K) && d >= K ? d + 1 : d];
    return r;
}

// The source of bin d in harmonic k's run.
template <int K>
__device__ __forceinline__ float src(const Run<K>& r, int d) { return r.v[(d * K + 8) >> 4]; }

// v[d] of the lane's 8 predecessors (ext[0 .. 8), from the left neighbour) and own bins
// (ext[8 .. 24)).
__device__ __forceinline__ float shift_in(float x) {
    // wave_shr:1 -- lane l receives lane l - 1's value (lane 0 keeps its own, unused)
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, x),
                                                                 __builtin_bit_cast(int, x), 0x138, 0xf, 0xf, false));
}

// Level h (width 2^h) of the lane's fundamentals j = B / 2^h + m, m < 16 / 2^h: the value
// of the reference's gather form (shaders/harmonic.comp:14-36): the best sum of the interval
// within [W, H) if it passes the threshold, else the sum at the interval's first bin in
// [W, H), 0 for an interval with none.  ext holds the sums of bins B - 8 .. B + 15, -inf
// outside [W, H).  EXACT: every bin of the wave's intervals lies in [W, H).
template <int LG, bool EXACT, bool SNAP>
__device__ __forceinline__ void level_out(const float* ext, int B, int W, int H, int F, float thr, uint32_t* res,
                                          uint32_t capacity, float* snap_row) {
    constexpr int width = 1 << LG, half = width / 2, count = 16 / width;
#pragma unroll
    for (int m = 0; m < count; ++m) {
        const int j = B / width + m;
        if (j >= F) break;
        const int lo = 8 + m * width - half;  // ext index of the interval's first bin
        float best = ext[lo];
#pragma unroll
        for (int d = 1; d < width; ++d) best = fmaxf(best, ext[lo + d]);
        float value;
        if constexpr (EXACT) {
            value = best > thr ? best : ext[lo];
        } else {  // clip the interval to [W, H)
            const int start = j * width - half;
            const int begin = max(W, start), end = min(H, start + width);
            float first = ext[lo];
#pragma unroll
            for (int d = 1; d < width; ++d)
                if (start + d == begin) first = ext[lo + d];
            value = begin >= end ? 0.0f : (best > thr ? best : first);
        }
        if constexpr (SNAP) snap_row[j] = value;
        if (j >= W && value > thr) append_candidate(res, capacity, uint32_t(j), uint32_t(LG), value);
    }
}

// One lane's 16 bins from B (lane 0: the 16 before the tile, which produce nothing).  s: the
// wave's first 16-bin block (stretch start / k), lb: the lane's block from s, m: the lane.
// The stretches are fetched two ahead of their use, in the order the sums need them.
template <bool EXACT, bool SNAP>
__device__ __forceinline__ void harmonic_lane(const float* __restrict__ P, const brp_params& p, int B, bool owner,
                                              uint32_t* res, float* snap, uint32_t snap_f, float* __restrict__ buf,
                                              int s, int lb, int m) {
    // The template's fields in registers: the LDS phases below are compiler memory barriers,
    // across which fields read from memory would be read again.
    const int W = int(p.window), H = int(p.harmonic), F = int(p.fundamental);
    const uint32_t cap = p.capacity;
    float thr[5];
#pragma unroll
    for (int h = 0; h < 5; ++h) thr[h] = p.threshold[h];
    const int Bc = max(B, 0);  // run addresses (tile 0's lane 0 reads from bin 0 on)
    float v[16], ext[24];
    // ext = left neighbour's last 8 sums ++ own 16, -inf outside [W, H) (EXACT: none are).
    auto extend = [&]() {
#pragma unroll
        for (int d = 0; d < 16; ++d) {
            if constexpr (EXACT) ext[8 + d] = v[d];
            else ext[8 + d] = (B + d >= W && B + d < H) ? v[d] : -INFINITY;
        }
#pragma unroll
        for (int d = 0; d < 8; ++d) ext[d] = shift_in(ext[16 + d]);
    };

    const Run<4> r4 = load_run<4>(P, Bc);
    const Run<2> r2 = load_run<2>(P, Bc);
    const Run<3> r3 = load_run<3>(P, Bc);
    const Run<1> r1 = load_run<1>(P, Bc);
    const Pieces<16> p16 = fetch_stretch<16>(P, s, m);
    const Pieces<8> p8 = fetch_stretch<8>(P, s, m);

    // Level 0 and the start of the sums: k = 16 (the bins themselves) and k = 8.
    {
        const Run<16> r16 = stage_run<16>(buf, p16, m, lb);
        if (owner && B < F) {
#pragma unroll
            for (int d = 0; d < 16; ++d) {
                const int i = B + d;
                if (i < F) {
                    if constexpr (SNAP) snap[i] = r16.v[d];
                    if (i >= W && r16.v[d] > thr[0]) append_candidate(res, cap, uint32_t(i), 0, r16.v[d]);
                }
            }
        }
        const Pieces<12> p12 = fetch_stretch<12>(P, s, m);
        const Run<8> r8 = stage_run<8>(buf, p8, m, lb);
        const Pieces<14> p14 = fetch_stretch<14>(P, s, m);
        // Historical order (shaders/hs_sum.glsl:2-10), at(k) = P[(i k + 8) >> 4].
#pragma unroll
        for (int d = 0; d < 16; ++d) v[d] = src(r16, d) + src(r8, d);
        extend();
        if (owner && B < 2 * F)
            level_out<1, EXACT, SNAP>(ext, B, W, H, F, thr[1], res, cap, snap + 1 * snap_f);
        {
            const Run<12> r12 = stage_run<12>(buf, p12, m, lb);
#pragma unroll
            for (int d = 0; d < 16; ++d) v[d] += src(r12, d) + src(r4, d);
        }
        const Pieces<10> p10 = fetch_stretch<10>(P, s, m);
        extend();
        if (owner && B < 4 * F)
            level_out<2, EXACT, SNAP>(ext, B, W, H, F, thr[2], res, cap, snap + 2 * snap_f);
        const Run<14> r14 = stage_run<14>(buf, p14, m, lb);
        const Pieces<6> p6 = fetch_stretch<6>(P, s, m);
        const Run<10> r10 = stage_run<10>(buf, p10, m, lb);
        const Pieces<15> p15 = fetch_stretch<15>(P, s, m);
        const Run<6> r6 = stage_run<6>(buf, p6, m, lb);
        const Pieces<13> p13 = fetch_stretch<13>(P, s, m);
#pragma unroll
        for (int d = 0; d < 16; ++d) v[d] += src(r14, d) + src(r10, d) + src(r6, d) + src(r2, d);
        extend();
        if (owner && B < 8 * F)
            level_out<3, EXACT, SNAP>(ext, B, W, H, F, thr[3], res, cap, snap + 3 * snap_f);
        float in[16];
        {
            const Run<15> r15 = stage_run<15>(buf, p15, m, lb);
            const Pieces<11> p11 = fetch_stretch<11>(P, s, m);
            const Run<13> r13 = stage_run<13>(buf, p13, m, lb);
            const Pieces<9> p9 = fetch_stretch<9>(P, s, m);
#pragma unroll
            for (int d = 0; d < 16; ++d) in[d] = src(r15, d) + src(r13, d);
            const Run<11> r11 = stage_run<11>(buf, p11, m, lb);
            const Pieces<7> p7 = fetch_stretch<7>(P, s, m);
            const Run<9> r9 = stage_run<9>(buf, p9, m, lb);
            const Pieces<5> p5 = fetch_stretch<5>(P, s, m);
#pragma unroll
            for (int d = 0; d < 16; ++d) in[d] = (in[d] + src(r11, d)) + src(r9, d);
            const Run<7> r7 = stage_run<7>(buf, p7, m, lb);
            const Run<5> r5 = stage_run<5>(buf, p5, m, lb);
#pragma unroll
            for (int d = 0; d < 16; ++d) v[d] += (((in[d] + src(r7, d)) + src(r5, d)) + src(r3, d)) + src(r1, d);
        }
    }
    extend();
    if (owner && B < 16 * F)
        level_out<4, EXACT, SNAP>(ext, B, W, H, F, thr[4], res, cap, snap + 4 * snap_f);
}

template <bool SNAP>
__global__ __launch_bounds__(64) void k_harmonic(FusedArgs a, FusedShape sh, uint32_t tiles) {
    const int lane = threadIdx.x;
    const uint32_t item = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t g = item / tiles, tile = item % tiles;
    const uint32_t slot = a.first + g;
    const brp_params& p = a.params[slot];
    const float* __restrict__ P = a.P + size_t(g) * sh.m;
    uint32_t* res = a.results + size_t(slot) * a.slot_words;
    float* snap = SNAP ? a.snap_summed + size_t(slot) * 5 * a.snap_f : nullptr;
    __shared__ float buf[kStageWords];
    const int i0 = int(tile) * kTileBins;
    const int B = i0 + (lane - 1) * kBinsPerLane;
    const int s = max(i0 / kBinsPerLane - 1, 0);
    const int lb = max(B, 0) / kBinsPerLane - s;
    // Every bin the wave's intervals touch (from 8 before the tile) in [W, H): no clipping.
    if (i0 - 8 >= int(p.window) && i0 + kTileBins <= int(p.harmonic))
        harmonic_lane<true, SNAP>(P, p, B, lane > 0, res, snap, a.snap_f, buf, s, lb, lane);
    else
        harmonic_lane<false, SNAP>(P, p, B, lane > 0, res, snap, a.snap_f, buf, s, lb, lane);
}

__global__ void k_pack_results(const uint32_t* __restrict__ results, uint32_t slot_words, uint32_t nslots,
                               uint32_t keep, uint32_t* __restrict__ staging) {
    const uint32_t slot = blockIdx.x;
    if (slot >= nslots) return;
    const uint4* src = reinterpret_cast<const uint4*>(results + size_t(slot) * slot_words);
    uint4* dst = reinterpret_cast<uint4*>(staging) + size_t(slot) * (keep + 1);
    const uint32_t count = results[size_t(slot) * slot_words];
    // The counter keeps counting past an overflow; never read beyond the slot's capacity.
    const uint32_t n = min(min(count, keep), slot_words / 4 - 1);
    for (uint32_t e = threadIdx.x; e <= n; e += blockDim.x) dst[e] = src[e];
}

// ------------------------------------------------------ resampled lengths --
// resampling_steps (src/science.cpp:41-45) of a template whose jump start is safe: the
// largest steps <= unpadded - 1 with float(steps) - offset(steps) < unpadded - 1 in the FP32
// rule, by the host's method: a scan down from 64 samples above the FP64 solution s* of
// s - offset(s) = unpadded - 1 (a fixed point with contraction factor tau omega < 1/2).  The
// host only defers templates for which no s more than 64 samples above s* can satisfy the
// rule, so the scan's result is the reference's; the device FP64 sine moves s* by far less
// than a sample.
__device__ uint32_t resampled_length(const brp_params& p, const float2* __restrict__ lut) {
    const float last = float(p.unpadded - 1);
    uint32_t steps = p.unpadded - 1;
    const double u1 = double(p.unpadded - 1);
    double s = u1;
    for (int it = 0; it < 200; ++it) {
        const double next =
            u1 + (double(p.tau) * sin_f64(double(p.omega) * (s * double(p.dt)) + double(p.psi)) * double(p.step_inv) -
                  double(p.s0));
        const bool done = fabs(next - s) < 1e-3;
        s = next;
        if (done) break;
    }
    const double start = floor(s) + 64.0;
    if (start < double(steps)) steps = start < 1.0 ? 1u : uint32_t(start);
    while (steps && float(steps) - sample_offset(steps, p, lut) >= last) --steps;
    return steps;
}

// The lengths of the templates the host left at steps = 0, in place: first the length
// parameters() computes with the double-sine S0 and plain arithmetic, whose emptiness is an
// error as in the reference (src/workload.cpp:110), then the template's own (flag-2
// arithmetic, single-sine S0).  error bit 1: a template produces an empty series.
__global__ __launch_bounds__(64) void k_search_steps(brp_params* params, uint32_t count,
                                                     const float2* __restrict__ lut, uint32_t* error) {
    const uint32_t t = blockIdx.x * 64 + threadIdx.x;
    if (t >= count) return;
    brp_params p = params[t];
    if (p.steps) return;
    brp_params plain = p;
    plain.flags = 0;
    plain.s0 = float(double(p.tau) * sin_f64(double(p.psi)) * double(p.step_inv));
    const uint32_t first = resampled_length(plain, lut);
    const uint32_t steps = resampled_length(p, lut);
    if (!first || !steps) atomicOr(error, 1u);
    params[t].steps = steps;
}

// Plain streaming copy, 16 bytes per lane per iteration (the counter calibration kernel).
__global__ __launch_bounds__(256) void brp_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) dst[i] = src[i];
}

// Pass-1 shapes compiled with the split-FP16 MFMA row FFT (the 2048- and 4096-point rows the
// 2^k and 3 * 2^k lengths use by default; both plans start with two radix-16 stages).
constexpr bool pass1_mfma_instance(int c, int s, int t) {
    return (c == 2048 && s == 2 && t == 256) || (c == 4096 && s == 2 && t == 512);
}

template <int C, int S, int T>
void launch1(hipStream_t st, const FusedShape& sh, const FusedArgs& a) {
    const dim3 grid(sh.blocks1 * a.count), block(T);
    const bool snap = a.snap_resampled != nullptr;
    // Index form per group (the same in validation runs with snapshots); FP64 phase runs
    // take the general form.
    if (a.fp64_phase) {
        if (snap) k_pass1<C, S, T, true, true, false, false, false><<<grid, block, 0, st>>>(a, sh);
        else k_pass1<C, S, T, false, true, false, false, false><<<grid, block, 0, st>>>(a, sh);
        return;
    }
    auto go = [&]<bool NN, bool FM, bool MF>() {
        if (snap) k_pass1<C, S, T, true, false, NN, FM, MF><<<grid, block, 0, st>>>(a, sh);
        else k_pass1<C, S, T, false, false, NN, FM, MF><<<grid, block, 0, st>>>(a, sh);
    };
    auto form = [&]<bool MF>() {
        if (a.nonneg && a.fma) go.template operator()<true, true, MF>();
        else if (a.nonneg) go.template operator()<true, false, MF>();
        else if (a.fma) go.template operator()<false, true, MF>();
        else go.template operator()<false, false, MF>();
    };
    if constexpr (pass1_mfma_instance(C, S, T)) {
        if (a.precision & 2u) {
            form.template operator()<true>();
            return;
        }
    }
    form.template operator()<false>();
}
template <int R, int S, int T>
void launch2(hipStream_t st, const FusedShape& sh, const FusedArgs& a) {
    if constexpr (plan_for(R, R * 2 * S / T).r[0] == 16) {
        if (a.precision & 1u) {
            if (a.snap_X) k_pass2<R, S, T, true, true><<<sh.blocks2 * a.count, T, 0, st>>>(a, sh);
            else k_pass2<R, S, T, false, true><<<sh.blocks2 * a.count, T, 0, st>>>(a, sh);
            return;
        }
    }
    if (a.snap_X) k_pass2<R, S, T, true, false><<<sh.blocks2 * a.count, T, 0, st>>>(a, sh);
    else k_pass2<R, S, T, false, false><<<sh.blocks2 * a.count, T, 0, st>>>(a, sh);
}

// The compiled kernel instances: f.template operator()<C, S1, T>() for pass 1 and
// <R, S2, T> for pass 2; false when (length, rows, threads) has no instance.
constexpr uint64_t inst(uint32_t len, uint32_t rows, uint32_t threads) {
    return (uint64_t(len) << 32) | (uint64_t(rows) << 16) | threads;
}
template <typename F>
bool with_pass1(uint32_t c, uint32_t s1, uint32_t t1, F&& f) {
    switch (inst(c, s1, t1)) {
#define BRP_P1(C, S, T) case inst(C, S, T): f.template operator()<C, S, T>(); return true;
        BRP_P1(64, 64, 256) BRP_P1(128, 32, 256) BRP_P1(256, 16, 256) BRP_P1(512, 16, 512)
        BRP_P1(1024, 16, 1024) BRP_P1(1024, 8, 512) BRP_P1(2048, 8, 1024) BRP_P1(2048, 4, 512) BRP_P1(2048, 2, 256) BRP_P1(2048, 2, 512)
        BRP_P1(4096, 4, 1024) BRP_P1(4096, 2, 512) BRP_P1(4096, 1, 256)
        BRP_P1(4096, 2, 1024) BRP_P1(4096, 1, 512)
        BRP_P1(384, 32, 512) BRP_P1(768, 16, 512) BRP_P1(1536, 8, 512) BRP_P1(3072, 2, 256)
#undef BRP_P1
        default: return false;
    }
}
template <typename F>
bool with_pass2(uint32_t r, uint32_t s2, uint32_t t2, F&& f) {
    switch (inst(r, s2, t2)) {
#define BRP_P2(R, S, T) case inst(R, S, T): f.template operator()<R, S, T>(); return true;
        BRP_P2(64, 32, 256) BRP_P2(128, 16, 256) BRP_P2(256, 16, 512) BRP_P2(256, 8, 256)
        BRP_P2(512, 16, 1024) BRP_P2(512, 8, 512) BRP_P2(512, 4, 256) BRP_P2(512, 4, 512)
        BRP_P2(1024, 8, 1024) BRP_P2(1024, 4, 512) BRP_P2(1024, 2, 256) BRP_P2(2048, 4, 1024) BRP_P2(2048, 2, 512)
        BRP_P2(192, 16, 256) BRP_P2(384, 8, 256) BRP_P2(768, 4, 256) BRP_P2(1536, 2, 256)
        BRP_P2(384, 16, 512) BRP_P2(768, 8, 512)
#undef BRP_P2
        default: return false;
    }
}

// Per series length M = N/2: pass-1 length C with S1 rows per block, pass-2 length R
// with S2 column pairs per block, and the threads of each block.  Blocks hold 16 elements per thread where the radix
// plan allows (up to 1024 threads, 16 waves per CU), 48 for the 3 * 2^k plans.  The first
// row of a length is the default; the others are the autotuner's alternatives (other
// LDS tiles C x S1 and 2 R x S2, and so other radix sequences).
struct ShapeRow { uint32_t m, c, r, s1, s2, t1, t2; };
constexpr ShapeRow kShapes[] = {
    {1u << 12, 64, 64, 64, 32, 256, 256},     {1u << 13, 128, 64, 32, 32, 256, 256},
    {1u << 14, 128, 128, 32, 16, 256, 256},   {1u << 15, 256, 128, 16, 16, 256, 256},
    {1u << 16, 256, 256, 16, 16, 256, 512},   {1u << 17, 512, 256, 16, 16, 512, 512},
    {1u << 18, 1024, 256, 16, 16, 1024, 512}, {1u << 19, 2048, 256, 2, 8, 256, 256},
    {1u << 20, 2048, 512, 2, 4, 256, 256},    {1u << 21, 2048, 1024, 2, 4, 256, 512},
    {1u << 22, 4096, 1024, 4, 8, 1024, 1024},
    // 3 * 2^k: the factor 3 in the pass-2 length R (24 elements per thread, radix 3, 8, ...),
    // so pass 1 keeps its power-of-two row FFT at 16 elements per thread
    {3u << 14, 256, 192, 16, 16, 256, 256},   {3u << 15, 256, 384, 16, 8, 256, 256},
    {3u << 16, 512, 384, 16, 8, 512, 256},    {3u << 17, 1024, 384, 16, 8, 1024, 256},
    {3u << 18, 2048, 384, 2, 8, 256, 256},    {3u << 19, 2048, 768, 2, 4, 256, 256},
    {3u << 20, 4096, 768, 2, 4, 512, 256},    {3u << 21, 4096, 1536, 2, 2, 512, 256},
    // alternatives
    {3u << 14, 384, 128, 32, 16, 512, 256},   {3u << 15, 768, 128, 16, 16, 512, 256},
    {3u << 16, 768, 256, 16, 16, 512, 512},   {3u << 17, 1536, 256, 8, 16, 512, 512},
    {3u << 18, 3072, 256, 2, 16, 256, 512},   {3u << 19, 3072, 512, 2, 16, 256, 1024},
    {3u << 20, 3072, 1024, 2, 8, 256, 1024},  {3u << 21, 3072, 2048, 2, 4, 256, 1024},
    {3u << 18, 4096, 192, 2, 16, 512, 256},   {3u << 19, 4096, 384, 2, 16, 512, 512},
    {3u << 20, 4096, 768, 2, 8, 512, 512},
    {1u << 21, 4096, 512, 4, 16, 1024, 1024},
    {1u << 19, 1024, 512, 16, 16, 1024, 1024}, {1u << 19, 4096, 128, 4, 16, 1024, 256},
    {1u << 20, 2048, 512, 8, 16, 1024, 1024},  {1u << 20, 1024, 1024, 16, 8, 1024, 1024},
    {1u << 21, 2048, 1024, 8, 8, 1024, 1024},  {1u << 21, 1024, 2048, 16, 4, 1024, 1024},
    {1u << 22, 2048, 2048, 8, 4, 1024, 1024},
    {3u << 18, 1536, 512, 8, 16, 512, 1024},   {3u << 18, 768, 1024, 16, 8, 512, 1024},
    {3u << 19, 1536, 1024, 8, 8, 512, 1024},   {3u << 19, 768, 2048, 16, 4, 512, 1024},
    {3u << 20, 1536, 2048, 8, 4, 512, 1024},
    // half-size blocks (512 threads, two blocks per CU)
    {1u << 19, 2048, 256, 4, 16, 512, 512},    {1u << 20, 4096, 256, 2, 16, 512, 512},
    {1u << 21, 4096, 512, 2, 8, 512, 512},     {1u << 21, 4096, 512, 2, 16, 512, 1024},
    {1u << 21, 4096, 512, 4, 8, 1024, 512},    {1u << 21, 2048, 1024, 4, 4, 512, 512},
    {1u << 22, 4096, 1024, 2, 4, 512, 512},
    // quarter-size blocks (256 threads, four blocks per CU)
    {1u << 21, 4096, 512, 1, 4, 256, 256},     {1u << 21, 4096, 512, 2, 4, 512, 256},
    {1u << 19, 2048, 256, 8, 16, 1024, 512},   {1u << 20, 4096, 256, 4, 16, 1024, 512},
    {3u << 18, 2048, 384, 4, 8, 512, 256},     {3u << 19, 4096, 384, 2, 8, 512, 256},
    {1u << 21, 4096, 512, 1, 8, 256, 512},     {1u << 21, 2048, 1024, 2, 2, 256, 256},
    {1u << 20, 4096, 256, 1, 8, 256, 256},
    // eight elements per thread (radix-8 plans, fewer registers, more waves per CU)
    {1u << 21, 4096, 512, 2, 4, 1024, 256},    {1u << 21, 4096, 512, 1, 4, 512, 256},
    {1u << 21, 4096, 512, 2, 4, 512, 512},     {1u << 21, 4096, 512, 1, 4, 512, 512},
    {1u << 21, 2048, 1024, 2, 4, 512, 512},
    // eight 1024-point rows per pass-1 block (64-byte pieces of each observation line)
    {1u << 21, 1024, 2048, 8, 2, 512, 512},
};

}  // namespace

namespace brp_engine {

bool fused_mfma_capable(const FusedShape& s, uint32_t precision) {
    const bool pass2 = plan_for(int(s.r), int(2 * s.r * s.s2 / s.t2)).r[0] == 16;
    return (!(precision & 1u) || pass2) && (!(precision & 2u) || pass1_mfma_instance(int(s.c), int(s.s1), int(s.t1)));
}

StageTable fft_stage_table(uint32_t len, uint32_t elements) {
    const Radices pl = plan_for(int(len), int(elements));
    StageTable tb;
    uint32_t ns = uint32_t(pl.r[0]);
    for (int st = 1; st < pl.count; ++st) {
        tb.offset[tb.stages] = tb.count;
        tb.ns[tb.stages] = ns;
        tb.radix[tb.stages] = uint32_t(pl.r[st]);
        tb.count += (uint32_t(pl.r[st]) - 1) * ns;
        ns *= uint32_t(pl.r[st]);
        ++tb.stages;
    }
    tb.len = len;
    return tb;
}

uint32_t fused_pass1_elements(const FusedShape& s) { return s.c * s.s1 / s.t1; }
uint32_t fused_pass2_elements(const FusedShape& s) { return 2 * s.r * s.s2 / s.t2; }

namespace {
// The shape of a table row, if both of its kernels are compiled.
bool shape_of(const ShapeRow& row, uint32_t n, FusedShape* out) {
    if (!with_pass1(row.c, row.s1, row.t1, []<int, int, int>() {}) ||
        !with_pass2(row.r, row.s2, row.t2, []<int, int, int>() {}))
        return false;
    FusedShape s;
    s.n = n; s.m = n / 2; s.c = row.c; s.r = row.r; s.s1 = row.s1; s.s2 = row.s2; s.t1 = row.t1; s.t2 = row.t2;
    s.blocks1 = row.r / row.s1;
    s.blocks2 = row.c / (2 * row.s2);
    s.mirror = row.s2 <= 4;
    s.n_shift = uint32_t(__builtin_ctz(n));
    s.n_odd = n >> s.n_shift;
    if (s.n_odd != 1 && s.n_odd != 3) return false;
    *out = s;
    return true;
}
}  // namespace

std::vector<FusedShape> fused_shapes(uint32_t n) {
    std::vector<FusedShape> out;
    for (const ShapeRow& row : kShapes) {
        FusedShape s;
        if (row.m == n / 2 && shape_of(row, n, &s)) out.push_back(s);
    }
    return out;
}

bool fused_supported(uint32_t n, uint32_t c, uint32_t s1, uint32_t s2, FusedShape* out, uint32_t t1, uint32_t t2) {
    const uint32_t m = n / 2;
    for (const ShapeRow& row : kShapes) {
        if (row.m != m) continue;
        if ((c && c != row.c) || (s1 && s1 != row.s1) || (s2 && s2 != row.s2) || (t1 && t1 != row.t1) ||
            (t2 && t2 != row.t2))
            continue;
        if (shape_of(row, n, out)) return true;
    }
    return false;
}

void fused_pass1(hipStream_t st, const FusedShape& s, const FusedArgs& a) {
    with_pass1(s.c, s.s1, s.t1, [&]<int C, int S, int T>() { launch1<C, S, T>(st, s, a); });
}

void fused_pass2(hipStream_t st, const FusedShape& s, const FusedArgs& a) {
    with_pass2(s.r, s.s2, s.t2, [&]<int R, int S, int T>() { launch2<R, S, T>(st, s, a); });
}

uint32_t harmonic_power_bins(uint32_t reach) {
    // Lane 63 of the last tile reads P[B .. B + 16) with B = tiles kTileBins - 16; the runs of
    // k < 16 end lower.
    return (reach + kTileBins - 1) / kTileBins * kTileBins;
}

void fused_harmonic(hipStream_t st, const FusedShape& s, const FusedArgs& a, uint32_t reach) {
    const uint32_t tiles = (reach + kTileBins - 1) / kTileBins;
    if (a.snap_summed) k_harmonic<true><<<tiles * a.count, 64, 0, st>>>(a, s, tiles);
    else k_harmonic<false><<<tiles * a.count, 64, 0, st>>>(a, s, tiles);
}

void search_steps(hipStream_t st, brp_params* params, uint32_t count, const float2* lut, uint32_t* error) {
    k_search_steps<<<(count + 63) / 64, 64, 0, st>>>(params, count, lut, error);
}

void copy16(hipStream_t st, const void* src, void* dst, size_t bytes) {
    brp_copy<<<2048, 256, 0, st>>>(static_cast<const uint4*>(src), static_cast<uint4*>(dst), bytes / 16);
}

void pack_results(hipStream_t st, const uint32_t* results, uint32_t slot_words, uint32_t nslots, uint32_t keep,
                  uint32_t* staging) {
    k_pack_results<<<nslots, 64, 0, st>>>(results, slot_words, nslots, keep, staging);
}

}  // namespace brp_engine
