// noalloc.hip -- brp_hip_search_noalloc: the fused pipeline (pass 1, pass 2, harmonic) for a
// few templates of one series with every intermediate buffer in a file-scope device array,
// for callers that cannot allocate device memory (code under HIP graph capture, serving loops
// that must not call hipMalloc).  The caller passes the series and the result slots, both in
// device memory, and a stream; the call only enqueues work on that stream.
//
// Nothing here allocates: the parameter blocks and the LUT travel as kernel arguments and are
// written into the scratch array by a kernel, the twiddle tables are computed on the device
// (double cos / sin of -2 pi t / len rounded once to float, the same rule the handle's host
// tables follow), and the slots are cleared by a kernel.  Every call on a GPU shares that
// GPU's scratch array: calls that overlap on one GPU race, and each call rewrites every part
// of the scratch it reads.
#include <mutex>
#include <string>

#include "device.h"
#include "engine.h"
#include "errors.h"
#include "lut.h"

namespace {

constexpr size_t kScratchBytes = size_t(BRP_NOALLOC_SCRATCH_BYTES);
constexpr int kMaxDevices = 64;
static_assert(BRP_NOALLOC_MAX_TEMPLATES <= 8, "parameter blocks travel as one kernel argument");

// The scratch array.  In host code `scratch` is a stand-in symbol, not a device address:
// host code only ever uses the per-device address from hipGetSymbolAddress.
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
unsigned char* g_scratch[kMaxDevices] = {};

struct Error {
    int code;
    std::string message;
};

void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error{BRP_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e)};
}

void need(bool ok, const std::string& message) {
    if (!ok) throw Error{BRP_E_INVALID, message};
}

// Address of the scratch array on the current device, looked up once per device with that
// device current.
unsigned char* scratch_here() {
    int dev = -1;
    hip_ok(hipGetDevice(&dev), "hipGetDevice");
    need(dev >= 0 && dev < kMaxDevices, "no-allocation search: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratch[dev];
}

// Parameter blocks and the LUT by value, written to the scratch by one block.
struct SetupArgs {
    brp_params params[BRP_NOALLOC_MAX_TEMPLATES];
    float2 lut[65];
    uint32_t count;
    brp_params* params_out;
    float2* lut_out;
};

__global__ void k_setup(SetupArgs s) {
    const uint32_t t = threadIdx.x;
    if (t < s.count) {
        s.params_out[t] = s.params[t];
        s.params_out[t].result_slot = t;
    }
    if (t < 65) s.lut_out[t] = s.lut[t];
}

// W_len^t for t = (i stride) mod len, i < count, of each table (or, for the FFT passes'
// per-stage tables, t = the stage table's exponent of entry i): double cos / sin of
// -2 pi t / len rounded once to float.
struct Table {
    float2* out;
    uint64_t len, count, stride;
    brp_engine::StageTable stage;  // stage.count > 0: a per-stage table
};
constexpr int kTables = 7;
struct TableArgs {
    Table table[kTables];
    uint64_t start[kTables + 1];  // prefix sums of the counts
};

__global__ void k_twiddles(TableArgs a) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.start[kTables]) return;
    int k = 0;
    while (i >= a.start[k + 1]) ++k;
    const Table& tb = a.table[k];
    const uint64_t e = i - a.start[k];
    const uint64_t t = (tb.stage.count ? tb.stage.exponent(uint32_t(e)) : e * tb.stride) % tb.len;
    const double ang = -6.283185307179586476925286766559 * double(t) / double(tb.len);
    tb.out[e] = make_float2(float(cos(ang)), float(sin(ang)));
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Carves the scratch array: returns the bytes used (may exceed the array; nothing is written
// through the pointers then).
struct Carve {
    unsigned char* base;
    size_t used = 0;
    template <class T>
    T* take(size_t count) {
        T* p = reinterpret_cast<T*>(base + used);
        used = align256(used + count * sizeof(T));
        return p;
    }
};

}  // namespace

extern "C" int brp_hip_search_noalloc(const brp_params* templates, size_t count, const float* input,
                                      size_t input_count, uint32_t* results, void* stream) {
    try {
        need(templates && input && results, "null argument");
        need(count >= 1 && count <= BRP_NOALLOC_MAX_TEMPLATES,
             "no-allocation search: 1 to " + std::to_string(BRP_NOALLOC_MAX_TEMPLATES) + " templates per call");
        const brp_params& p0 = templates[0];
        for (size_t i = 0; i < count; ++i) {
            const brp_params& p = templates[i];
            need(p.n == p0.n && p.unpadded == p0.unpadded && p.capacity == p0.capacity &&
                     (p.flags & 6u) == (p0.flags & 6u),
                 "no-allocation search: every template must share n, unpadded, capacity and arithmetic flags");
            need(p.steps && p.steps <= p.unpadded && p.unpadded <= p.n && p.capacity,
                 "no-allocation search: invalid template");
            need(p.window <= p.fundamental && p.harmonic <= p.n / 2 && p.fundamental <= p.n / 2,
                 "template search range outside the spectrum");
        }
        // The pass-1 gather reads 12 bytes from the last sample on: two floats past the series
        // (loaded, never used).
        need(input_count >= size_t(p0.unpadded) + 2,
             "no-allocation search: the input must hold unpadded + 2 readable floats");
        brp_engine::FusedShape sh;
        need(brp_engine::fused_supported(p0.n, 0, 0, 0, &sh),
             "fused engine: unsupported series length (N = 2^13..2^23 or 3*2^15..3*2^22)");

        // Lay out the scratch: parameters, LUT, twiddle tables, partial sums, means, the
        // exchange and the power (with the harmonic kernel's 2048-float tail).
        unsigned char* base = scratch_here();
        Carve cv{base};
        brp_params* params = cv.take<brp_params>(count);
        float2* lut = cv.take<float2>(65);
        TableArgs ta{};
        const brp_engine::StageTable st1 = brp_engine::fft_stage_table(sh.c, brp_engine::fused_pass1_elements(sh));
        const brp_engine::StageTable st2 = brp_engine::fft_stage_table(sh.r, brp_engine::fused_pass2_elements(sh));
        const uint64_t lens[kTables] = {sh.c, sh.r, sh.m, sh.m, 16ull * sh.r, sh.n, sh.n};
        const uint64_t counts[kTables] = {st1.count, st2.count, 2048, (sh.m + 2047) / 2048, 16ull * sh.r, 2048,
                                          (sh.n + 2047) / 2048};
        const uint64_t strides[kTables] = {1, 1, 1, 2048, 1, 1, 2048};
        for (int k = 0; k < kTables; ++k) {
            ta.table[k] = Table{cv.take<float2>(counts[k]), lens[k], counts[k], strides[k], {}};
            ta.start[k + 1] = ta.start[k] + counts[k];
        }
        ta.table[0].stage = st1;
        ta.table[1].stage = st2;
        double* partial = cv.take<double>(count * sh.blocks1);
        float* mean = cv.take<float>(count);
        float2* Y = cv.take<float2>(count * size_t(sh.m));
        float* P = cv.take<float>(count * size_t(sh.m) + 2048);
        if (cv.used > kScratchBytes)
            throw Error{BRP_E_INVALID, "no-allocation search: N = " + std::to_string(p0.n) + " x " +
                                           std::to_string(count) + " templates needs " + std::to_string(cv.used) +
                                           " bytes of scratch, more than its " + std::to_string(kScratchBytes)};

        const hipStream_t st = static_cast<hipStream_t>(stream);
        (void)hipGetLastError();
        SetupArgs sa{};
        for (size_t i = 0; i < count; ++i) sa.params[i] = templates[i];
        for (int i = 0; i < 65; ++i) sa.lut[i] = make_float2(brp_lut::kTable[2 * i], brp_lut::kTable[2 * i + 1]);
        sa.count = uint32_t(count);
        sa.params_out = params;
        sa.lut_out = lut;
        k_setup<<<1, 128, 0, st>>>(sa);
        k_twiddles<<<uint32_t((ta.start[kTables] + 255) / 256), 256, 0, st>>>(ta);

        const uint32_t slot_words = 4 + 4 * p0.capacity;
        brp_engine::plain_clear_slots(st, results, slot_words, uint32_t(count));

        brp_engine::FusedArgs a;
        a.params = params;
        a.x = input;
        a.lut = lut;
        a.tw_c = ta.table[0].out;
        a.tw_r = ta.table[1].out;
        a.twm_lo = ta.table[2].out;
        a.twm_hi = ta.table[3].out;
        a.tw_rq = ta.table[4].out;
        a.twn_lo = ta.table[5].out;
        a.twn_hi = ta.table[6].out;
        a.Y = Y;
        a.P = P;
        a.partial = partial;
        a.mean = mean;
        a.results = results;
        a.slot_words = slot_words;
        a.first = 0;
        a.count = uint32_t(count);
        a.fma = (p0.flags & 2u) ? 1u : 0u;
        bool nonneg = true;
        uint64_t reach = 0;
        for (size_t i = 0; i < count; ++i) {
            const brp_params& p = templates[i];
            nonneg = nonneg && p.omega >= 0.0f && p.psi >= 0.0f && p.dt >= 0.0f;
            reach = std::max<uint64_t>(reach, std::min<uint64_t>(16ull * p.fundamental, uint64_t(p.harmonic) + 8));
        }
        a.nonneg = nonneg ? 1u : 0u;
        a.power_bins = brp_engine::harmonic_power_bins(uint32_t(reach));
        brp_engine::fused_pass1(st, sh, a);
        brp_engine::fused_pass2(st, sh, a);
        brp_engine::fused_harmonic(st, sh, a, uint32_t(reach));
        hip_ok(hipGetLastError(), "no-allocation search launch");
        return BRP_OK;
    } catch (const Error& e) {
        return brp::fail(e.code, e.message);
    } catch (const std::exception& e) {
        return brp::fail(BRP_E_RUNTIME, e.what());
    }
}
