// engine.hip -- the brp_hip handle: device memory, tables, launch sequences, result
// readback and the bank loop behind the C ABI of include/brp_hip.h.  It takes the place
// of the reference's GPU::Impl (src/vulkan.cpp:21-277) and GPU methods
// (src/vulkan.cpp:278-406), and of the bank loop of src/workunit_main.cpp:77-136.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <cxxabi.h>

#include <map>

#include "brp_hip.h"
#include "counters.h"
#include "engine.h"
#include "errors.h"
#include "lut.h"
#include "toplist_state.h"

namespace {

thread_local std::string g_error;

struct Failure : std::runtime_error {
    int code;
    Failure(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Failure(BRP_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPC(expr) hip_check((expr), #expr)

void require(bool ok, const char* message) {
    if (!ok) throw Failure(BRP_E_INVALID, message);
}

double process_cpu_ms() {
    timespec t{};
    clock_gettime(CLOCK_PROCESS_CPUTIME_ID, &t);
    return t.tv_sec * 1000.0 + t.tv_nsec / 1e6;
}

bool pow2(uint32_t v) { return v && !(v & (v - 1)); }

constexpr double kTwoPi = 6.283185307179586476925286766559;

// W_len^t = e^{-2 pi i t / len} for t in [first, first + count) step `stride`,
// evaluated in double and rounded once to FP32.
std::vector<float2> twiddles(uint64_t len, uint64_t count, uint64_t stride = 1) {
    std::vector<float2> w(count);
    for (uint64_t i = 0; i < count; ++i) {
        const uint64_t t = (i * stride) % len;
        const double a = -kTwoPi * double(t) / double(len);
        w[i] = make_float2(float(std::cos(a)), float(std::sin(a)));
    }
    return w;
}

// The per-stage twiddle table of a fused FFT pass, each entry rounded as twiddles() rounds.
std::vector<float2> stage_twiddles(const brp_engine::StageTable& tb) {
    std::vector<float2> w(tb.count);
    for (uint32_t e = 0; e < tb.count; ++e) {
        const double a = -kTwoPi * double(tb.exponent(e) % tb.len) / double(tb.len);
        w[e] = make_float2(float(std::cos(a)), float(std::sin(a)));
    }
    return w;
}

template <class F>
int guarded(F&& body) {
    try {
        body();
        return BRP_OK;
    } catch (const Failure& f) {
        g_error = f.what();
        return f.code;
    } catch (const std::exception& e) {
        g_error = e.what();
        return BRP_E_RUNTIME;
    }
}

}  // namespace

namespace brp {
int fail(int code, const std::string& message) {
    g_error = message;
    return code;
}
}  // namespace brp

struct brp_hip {
    brp_params base{};
    brp_options opt{};
    hipStream_t stream = nullptr;
    std::vector<std::pair<void*, size_t>> allocations;
    uint64_t allocated = 0;
    // Two halves of slots, parameters, staging and graph: the public calls use half 0; the
    // bank loop alternates halves so that one batch runs while the previous is retained.
    std::vector<brp_params> templates[2];

    float* x = nullptr;
    float2* lut = nullptr;
    uint32_t* results = nullptr;
    uint32_t slot_words = 0;
    uint32_t keep = 64;
    uint32_t* staging_dev = nullptr;
    uint32_t* staging_host = nullptr;
    brp_params* params_dev = nullptr;
    brp_params* params_host = nullptr;

    // plain engine
    brp_engine::PlainBuffers pb;
    float* power_buf = nullptr;  // n/2 power + 5 F summed (the reference's buffer 4 layout)

    // fused engine: one record per observation (series); templates pick theirs by
    // (n, unpadded).  Exchange / power / partial buffers are shared, sized for the
    // largest series.
    struct Series {
        uint32_t n = 0, unpadded = 0;
        brp_engine::FusedShape shape;
        brp_engine::FusedArgs tables;  // x, lut and the twiddle tables of this length
    };
    std::vector<Series> series;
    brp_engine::FusedArgs fa;        // shared buffers and settings
    uint32_t group = 1;
    uint32_t lanes = 1;              // groups in flight, on stream and side
    bool timed = true;               // per-kernel events
    uint32_t stride = 1;             // on every stride-th launch group
    hipStream_t side = nullptr;      // second lane
    hipEvent_t fork = nullptr, join = nullptr;
    size_t lane_y = 0, lane_p = 0, lane_partial = 0;  // per-lane buffer strides (elements)
    uint32_t m_max = 0, blocks1_max = 0, n_max = 0;   // what the shared buffers hold
    // A launch group: `count` consecutive slots from `first`, one series, one index form.
    struct Group {
        uint32_t first, count, series, nonneg, reach;  // reach: harmonic bins to cover
        bool operator==(const Group& o) const {
            return first == o.first && count == o.count && series == o.series && nonneg == o.nonneg &&
                   reach == o.reach;
        }
    };
    std::vector<Group> groups[2];    // of the last launch of each half
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
    std::vector<Group> graph_groups[2];  // what each captured graph runs
    bool graph_events = true;

    std::vector<hipEvent_t> events[2];
    hipEvent_t done[2] = {nullptr, nullptr};
    std::vector<double> stage_ms_per_template[2];  // 4 per template of the last run
    double kernel_ms[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    bool ran[2] = {false, false};
    double submit_ms = 0;
    double total_ms[4] = {0, 0, 0, 0};  // since the last brp_hip_kernel_ms
    uint64_t total_launches = 0;        // launch sequences timed since then
    uint64_t total_templates = 0;       // templates those covered
    uint32_t timed_launches[2] = {0, 0}, timed_templates[2] = {0, 0};  // of the last run of each half

    // Device toplist (brp_hip_bank_search_device): while `device_top` is set, launch()
    // raises each slot's thresholds before the kernels and runs the retain after them on
    // top_stream, where it overlaps the next batch.  A batch's thresholds come from the last
    // rows the retain of the previous batch on the same half recorded (dtop_lowest[half]),
    // as the host loop freezes them, so they do not depend on timing.  Per slot: its table
    // and its orbit (period, tau, psi as float).
    bool device_top = false;
    uint32_t dtop_rows = 0, dtop_tables = 0, dtop_capacity = 0;  // capacity: tables allocated
    brp_engine::TopRow* dtop = nullptr;
    float* dtop_lowest[2] = {nullptr, nullptr};  // per half: tables x 5
    uint32_t* dtop_error = nullptr;
    hipStream_t top_stream = nullptr;
    hipEvent_t kernels_done[2] = {nullptr, nullptr}, retained[2] = {nullptr, nullptr};
    bool retain_pending[2] = {false, false};
    uint32_t* slot_table_dev = nullptr;
    float4* slot_orbit_dev = nullptr;
    uint32_t* slot_table_host = nullptr;  // pinned, 2 * batch
    float4* slot_orbit_host = nullptr;    // pinned, 2 * batch

    // Templates whose resampled length the device computes (brp_hip_bank_search on the fused
    // engine leaves their steps at 0): per half, whether the batch has any, and the error
    // word of the length kernel (device and pinned host copies).
    bool deferred[2] = {false, false};
    uint32_t* steps_error_dev = nullptr;
    uint32_t* steps_error_host = nullptr;

    size_t slot_base(int half) const { return size_t(half) * opt.batch; }

    template <class T>
    T* alloc(size_t count) {
        void* p = nullptr;
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        HIPC(hipMalloc(&p, bytes));
        allocations.emplace_back(p, bytes);
        allocated += bytes;
        return static_cast<T*>(p);
    }
    void release(void* p) {
        if (!p) return;
        for (auto it = allocations.begin(); it != allocations.end(); ++it)
            if (it->first == p) {
                HIPC(hipFree(p));
                allocated -= it->second;
                allocations.erase(it);
                return;
            }
    }
    template <class T>
    T* upload(const std::vector<T>& v) {
        T* p = alloc<T>(v.size());
        HIPC(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return p;
    }
    hipEvent_t event(int half, size_t i) {
        while (events[half].size() <= i) {
            hipEvent_t e;
            HIPC(hipEventCreate(&e));
            events[half].push_back(e);
        }
        return events[half][i];
    }
    ~brp_hip() {
        if (stream) hipStreamSynchronize(stream);
        if (side) hipStreamSynchronize(side);
        if (top_stream) hipStreamSynchronize(top_stream);
        for (auto g : graph_exec) if (g) hipGraphExecDestroy(g);
        for (auto& list : events) for (hipEvent_t e : list) hipEventDestroy(e);
        for (hipEvent_t e : done) if (e) hipEventDestroy(e);
        for (auto& a : allocations) hipFree(a.first);
        if (staging_host) hipHostFree(staging_host);
        if (params_host) hipHostFree(params_host);
        if (slot_table_host) hipHostFree(slot_table_host);
        if (slot_orbit_host) hipHostFree(slot_orbit_host);
        if (steps_error_host) hipHostFree(steps_error_host);
        for (hipEvent_t e : kernels_done) if (e) hipEventDestroy(e);
        for (hipEvent_t e : retained) if (e) hipEventDestroy(e);
        if (top_stream) hipStreamDestroy(top_stream);
        if (fork) hipEventDestroy(fork);
        if (join) hipEventDestroy(join);
        if (side) hipStreamDestroy(side);
        if (stream) hipStreamDestroy(stream);
    }
};

namespace {

void check_options(const brp_params& p, const brp_options& o) {
    require(o.engine == BRP_ENGINE_PLAIN || o.engine == BRP_ENGINE_FUSED, "unknown engine");
    // src/vulkan.cpp:279-283
    require(!o.fused_power || o.packed, "fused power requires packed real FFT");
    require(!o.radix8_store || o.radix8, "paired-store radix 8 requires radix 8");
    require(!o.radix8 || (o.stockham && !o.tile), "radix8 requires Stockham without a shared tile");
    // src/vulkan.cpp:116-122
    const uint32_t fft_n = o.packed ? p.n / 2 : p.n;
    const uint32_t fft_base = pow2(fft_n) ? fft_n : fft_n / 3;
    require(o.tile <= fft_base, "FFT tile exceeds radix-2 block");
    require(!(o.stockham && o.tile), "Stockham requires --tile 0");
    require(o.workgroup && o.workgroup <= 256 && o.batch && o.batch <= 4096 &&
                (!o.tile || (pow2(o.tile) && o.tile >= 64 && o.tile <= 4096 && o.tile <= p.n)),
            "invalid workgroup, FFT tile, or batch");
    require(o.precision <= 3, "unknown precision variant");
    require(!o.precision || o.engine == BRP_ENGINE_FUSED, "the split-FP16 precision variant needs the fused engine");
    require(!(o.precision & 2u) || !o.fp64_phase, "the split-FP16 pass-1 variant runs the FP32 index rule only");
}

void check_template(const brp_params& base, const brp_params& next) {
    // src/vulkan.cpp:270-276
    require((next.flags & 6) == (base.flags & 6), "arithmetic specialization changed; rebuild pipelines");
    require(next.n == base.n && next.unpadded == base.unpadded && next.fundamental == base.fundamental &&
                next.harmonic == base.harmonic && next.capacity == base.capacity && next.steps &&
                next.steps <= next.unpadded,
            "template dimensions changed; rebuild resources");
    require(next.window <= next.fundamental && next.harmonic <= next.n / 2 && next.fundamental <= next.n / 2,
            "template search range outside the spectrum");
}

void create_plain(brp_hip& h) {
    const brp_params& p = h.base;
    auto& b = h.pb;
    b.n = p.n;
    b.packed = h.opt.packed ? 1u : 0u;
    b.fft_n = b.packed ? p.n / 2 : p.n;
    b.x = h.x;
    b.lut = h.lut;
    b.raw = h.alloc<float>(p.unpadded);
    b.index = h.opt.snapshot ? h.alloc<int32_t>(p.unpadded) : nullptr;
    b.partial = h.alloc<double>((p.unpadded + 255) / 256);
    b.mean = h.alloc<float>(1);
    b.a = h.alloc<float2>(b.fft_n + 1);
    b.scratch = h.alloc<float2>(b.fft_n + 1);
    b.X = h.alloc<float2>(p.n / 2 + 1);
    b.tw_fft = h.upload(twiddles(b.fft_n, b.fft_n));
    b.tw_real = h.upload(twiddles(p.n, p.n / 2 + 1));
    h.power_buf = h.alloc<float>(size_t(p.n / 2) + 5ull * p.fundamental);
    HIPC(hipMemset(h.power_buf, 0, sizeof(float) * (size_t(p.n / 2) + 5ull * p.fundamental)));
}

// Invalidates the captured graphs (they hold buffer addresses and launch plans).
void drop_graphs(brp_hip& h) {
    for (int half = 0; half < 2; ++half) {
        if (h.graph_exec[half]) HIPC(hipGraphExecDestroy(h.graph_exec[half]));
        h.graph_exec[half] = nullptr;
        h.graph_groups[half].clear();
    }
}

// (Re)allocates the shared fused buffers for the largest series: per lane `group` slots
// of exchange (M complex), power (M floats) and pass-1 partial sums; snapshots per batch
// slot at the largest N.
void size_fused_buffers(brp_hip& h) {
    uint32_t m = 0, b1 = 0, n = 0;
    for (const auto& s : h.series) {
        m = std::max(m, s.shape.m);
        b1 = std::max(b1, s.shape.blocks1);
        n = std::max(n, s.n);
    }
    if (m <= h.m_max && b1 <= h.blocks1_max && n <= h.n_max) return;
    drop_graphs(h);
    auto& a = h.fa;
    h.release(a.Y);
    h.release(a.P);
    h.release(a.partial);
    h.m_max = m;
    h.blocks1_max = b1;
    h.n_max = n;
    h.lane_y = size_t(h.group) * m;
    h.lane_p = size_t(h.group) * m;
    h.lane_partial = size_t(h.group) * b1;
    a.Y = h.alloc<float2>(h.lanes * h.lane_y);
    // 2048 floats past the last slot: the harmonic kernel reads the dyadic-sum sources of
    // its last tile's bins unconditionally (up to H + 1031 < M + 2048) and masks them.
    a.P = h.alloc<float>(h.lanes * h.lane_p + 2048);
    a.partial = h.alloc<double>(h.lanes * h.lane_partial);
    if (h.opt.snapshot) {
        h.release(a.snap_resampled);
        h.release(a.snap_index);
        h.release(a.snap_X);
        h.release(a.snap_summed);
        a.snap_n = n;
        a.snap_m1 = m + 1;
        a.snap_f = n / 2;
        a.snap_resampled = h.alloc<float>(size_t(h.opt.batch) * a.snap_n);
        a.snap_index = h.alloc<int32_t>(size_t(h.opt.batch) * a.snap_n);
        a.snap_X = h.alloc<float2>(size_t(h.opt.batch) * a.snap_m1);
        a.snap_summed = h.alloc<float>(size_t(h.opt.batch) * 5 * a.snap_f);
    }
}

// Registers an observation of n padded / unpadded samples already on the device (with
// four zero guard samples after it): its four-step shape (the option overrides where
// they fit this length, else the default) and twiddle tables.
void add_fused_series(brp_hip& h, const brp_params& p, float* x) {
    for (const auto& s : h.series)
        require(!(s.n == p.n && s.unpadded == p.unpadded), "a series of this length is already registered");
    brp_hip::Series s;
    s.n = p.n;
    s.unpadded = p.unpadded;
    // The first compiled shape of this length that fits the overrides (and, for the
    // split-FP16 variant, that has the radix-16 stages it runs on the matrix cores); added
    // series fall back to
    // ignoring the overrides.
    const auto pick = [&](bool overrides) {
        for (const auto& cand : brp_engine::fused_shapes(p.n)) {
            if (overrides && ((h.opt.split_c && cand.c != h.opt.split_c) ||
                              (h.opt.split_s1 && cand.s1 != h.opt.split_s1) ||
                              (h.opt.split_s2 && cand.s2 != h.opt.split_s2) ||
                              (h.opt.split_t1 && cand.t1 != h.opt.split_t1) ||
                              (h.opt.split_t2 && cand.t2 != h.opt.split_t2)))
                continue;
            if (h.opt.precision && !brp_engine::fused_mfma_capable(cand, h.opt.precision)) continue;
            s.shape = cand;
            return true;
        }
        return false;
    };
    if (!pick(true))
        require(!h.series.empty() && pick(false),
                h.opt.precision ? "split-FP16 variant: no shape of this length has the radix-16 stages it asks for"
                                : "fused engine: unsupported series length (N = 2^13..2^23 or 3*2^15..3*2^22) or split");
    const auto& sh = s.shape;
    auto& a = s.tables;
    a.x = x;
    a.lut = h.lut;
    a.tw_c = h.upload(stage_twiddles(brp_engine::fft_stage_table(sh.c, brp_engine::fused_pass1_elements(sh))));
    a.tw_r = h.upload(stage_twiddles(brp_engine::fft_stage_table(sh.r, brp_engine::fused_pass2_elements(sh))));
    a.twm_lo = h.upload(twiddles(sh.m, 2048));
    a.twm_hi = h.upload(twiddles(sh.m, (sh.m + 2047) / 2048, 2048));
    a.tw_rq = h.upload(twiddles(16ull * sh.r, 16ull * sh.r));
    a.twn_lo = h.upload(twiddles(sh.n, 2048));
    a.twn_hi = h.upload(twiddles(sh.n, (sh.n + 2047) / 2048, 2048));
    h.series.push_back(s);
    size_fused_buffers(h);
}

void create_fused(brp_hip& h) {
    h.group = h.opt.snapshot ? h.opt.batch : std::min<uint32_t>(h.opt.batch, h.opt.group ? h.opt.group : 8);
    h.lanes = h.opt.snapshot ? 1 : (h.opt.lanes ? h.opt.lanes : 1);
    require(h.lanes <= 2, "fused engine: lanes must be 1 or 2");
    h.timed = !h.opt.untimed;
    h.stride = h.opt.timing_stride ? h.opt.timing_stride : 1;
    auto& a = h.fa;
    a.mean = h.alloc<float>(2 * size_t(h.opt.batch));
    h.steps_error_dev = h.alloc<uint32_t>(2);
    HIPC(hipHostMalloc(reinterpret_cast<void**>(&h.steps_error_host), 2 * sizeof(uint32_t), hipHostMallocDefault));
    h.steps_error_host[0] = h.steps_error_host[1] = 0;
    a.results = h.results;
    a.slot_words = h.slot_words;
    a.params = h.params_dev;
    a.fp64_phase = h.opt.fp64_phase;
    a.precision = h.opt.precision;
    if (h.lanes == 2) {
        HIPC(hipStreamCreateWithFlags(&h.side, hipStreamNonBlocking));
        HIPC(hipEventCreateWithFlags(&h.fork, hipEventDisableTiming));
        HIPC(hipEventCreateWithFlags(&h.join, hipEventDisableTiming));
    }
    add_fused_series(h, h.base, h.x);
}

// Series of a template: the registered observation with its (n, unpadded).
uint32_t series_of(const brp_hip& h, const brp_params& p) {
    for (uint32_t i = 0; i < h.series.size(); ++i)
        if (h.series[i].n == p.n && h.series[i].unpadded == p.unpadded) return i;
    throw Failure(BRP_E_INVALID, "template dimensions changed; rebuild resources (no series of this length)");
}

// The fixed-dimension rule of src/vulkan.cpp:270-276, lifted for the fused engine: a
// template may name any registered series (its n and unpadded), with its own F and H;
// the result capacity and the arithmetic flags stay those of the handle.
void check_template(const brp_hip& h, const brp_params& next, bool steps_deferred = false) {
    if (h.opt.engine == BRP_ENGINE_PLAIN) return check_template(h.base, next);
    require((next.flags & 6) == (h.base.flags & 6), "arithmetic specialization changed; rebuild pipelines");
    (void)series_of(h, next);
    require(next.capacity == h.base.capacity && (next.steps || steps_deferred) && next.steps <= next.unpadded,
            "template dimensions changed; rebuild resources");
    require(next.window <= next.fundamental && next.harmonic <= next.n / 2 && next.fundamental <= next.n / 2,
            "template search range outside the spectrum");
}

// The plain launch sequence for slot r (src/vulkan.cpp:221-266), with events at the
// four stage boundaries.
void enqueue_plain(brp_hip& h, uint32_t r) {
    const brp_params& p = h.templates[0][r];
    hipStream_t st = h.stream;
    uint32_t* slot = h.results + size_t(r) * h.slot_words;
    brp_engine::plain_clear_slots(st, slot, h.slot_words, 1);
    HIPC(hipEventRecord(h.event(0, 5 * r), st));
    brp_engine::plain_resample(st, h.pb, p, slot);
    HIPC(hipEventRecord(h.event(0, 5 * r + 1), st));
    const bool fused_power = h.opt.fused_power != 0;
    float2* z = brp_engine::plain_fft(st, h.pb, h.opt.radix8 != 0, h.opt.radix4 != 0 || h.opt.radix8 != 0);
    brp_engine::plain_spectrum(st, h.pb, z, fused_power, h.power_buf);
    HIPC(hipEventRecord(h.event(0, 5 * r + 2), st));
    if (!fused_power || !h.pb.packed) brp_engine::plain_power(st, h.pb, h.power_buf);
    HIPC(hipEventRecord(h.event(0, 5 * r + 3), st));
    brp_engine::plain_harmonic(st, h.power_buf, h.power_buf + p.n / 2, p, h.opt.scatter != 0, slot);
    HIPC(hipEventRecord(h.event(0, 5 * r + 4), st));
}

// The launch groups of `half`: runs of consecutive slots of one series, cut into at
// most `group` templates.  Per group: whether every template's orbital phase is
// non-negative for all samples (omega, psi, dt >= 0, which selects the shorter exact
// index form) and the harmonic tiles that cover every fundamental's rounding interval
// (centres j 2^h < 16 F, intervals that start below H).
std::vector<brp_hip::Group> plan_groups(const brp_hip& h, int half) {
    std::vector<brp_hip::Group> out;
    const auto& tp = h.templates[half];
    const uint32_t count = uint32_t(tp.size());
    uint32_t first = 0;
    while (first < count) {
        const uint32_t s = series_of(h, tp[first]);
        uint32_t end = first + 1;
        while (end < count && end - first < h.group && series_of(h, tp[end]) == s) ++end;
        bool nn = true;
        uint64_t reach = 0;
        for (uint32_t i = first; i < end; ++i) {
            nn = nn && tp[i].omega >= 0.0f && tp[i].psi >= 0.0f && tp[i].dt >= 0.0f;
            reach = std::max(reach, std::min<uint64_t>(16ull * tp[i].fundamental, uint64_t(tp[i].harmonic) + 8));
        }
        out.push_back({first, end - first, s, nn ? 1u : 0u, uint32_t(reach)});
        first = end;
    }
    return out;
}

// The fused sequence for the templates of `half`: every launch group in three launches,
// then the packing of the results for one readback.  With two lanes, even groups run on
// the handle's stream and odd groups on the side stream, each lane with its own
// exchange / power / partial buffers, so one group's launches overlap the other's (and
// the gaps between them).
void enqueue_fused(brp_hip& h, int half, bool capturing) {
    const auto& groups = h.groups[half];
    const uint32_t count = uint32_t(h.templates[half].size());
    const size_t base = h.slot_base(half);
    uint32_t* results = h.results + base * h.slot_words;
    brp_engine::plain_clear_slots(h.stream, results, h.slot_words, count);
    const bool ev = h.timed && (!capturing || h.graph_events);
    if (h.lanes == 2) {
        HIPC(hipEventRecord(h.fork, h.stream));
        HIPC(hipStreamWaitEvent(h.side, h.fork, 0));
    }
    hipStream_t st = h.stream;
    uint32_t gi_now = 0;
    auto record = [&](size_t i) {
        if (!ev || gi_now % h.stride) return;
        if (capturing) HIPC(hipEventRecordWithFlags(h.event(half, i), st, hipEventRecordExternal));
        else HIPC(hipEventRecord(h.event(half, i), st));
    };
    for (uint32_t gi = 0; gi < groups.size(); ++gi) {
        gi_now = gi;
        const auto& grp = groups[gi];
        const auto& ser = h.series[grp.series];
        brp_engine::FusedArgs a = h.fa;
        const auto& tb = ser.tables;
        a.x = tb.x;
        a.lut = tb.lut;
        a.tw_c = tb.tw_c;
        a.tw_r = tb.tw_r;
        a.twm_lo = tb.twm_lo;
        a.twm_hi = tb.twm_hi;
        a.tw_rq = tb.tw_rq;
        a.twn_lo = tb.twn_lo;
        a.twn_hi = tb.twn_hi;
        const uint32_t lane = gi % h.lanes;
        st = lane ? h.side : h.stream;
        a.Y += lane * h.lane_y;
        a.P += lane * h.lane_p;
        a.partial += lane * h.lane_partial;
        if (h.opt.snapshot) a.P += size_t(grp.first) * h.m_max;  // snapshots keep every slot's power
        a.first = uint32_t(base) + grp.first;
        a.count = grp.count;
        a.fma = (h.base.flags & 2u) ? 1u : 0u;
        a.nonneg = grp.nonneg;
        a.power_bins = h.opt.snapshot ? 0u : brp_engine::harmonic_power_bins(grp.reach);
        record(4 * gi);
        brp_engine::fused_pass1(st, ser.shape, a);
        record(4 * gi + 1);
        brp_engine::fused_pass2(st, ser.shape, a);
        record(4 * gi + 2);
        brp_engine::fused_harmonic(st, ser.shape, a, grp.reach);
        record(4 * gi + 3);
    }
    if (h.lanes == 2) {
        HIPC(hipEventRecord(h.join, h.side));
        HIPC(hipStreamWaitEvent(h.stream, h.join, 0));
    }
    st = h.stream;
    brp_engine::pack_results(st, results, h.slot_words, count, h.keep, h.staging_dev + base * (h.keep + 1) * 4);
}

void run_fused(brp_hip& h, int half) {
    h.groups[half] = plan_groups(h, half);
    if (!h.opt.graph || h.opt.snapshot) {
        enqueue_fused(h, half, false);
        return;
    }
    if (!h.graph_exec[half] || h.graph_groups[half] != h.groups[half]) {
        if (h.graph_exec[half]) { HIPC(hipGraphExecDestroy(h.graph_exec[half])); h.graph_exec[half] = nullptr; }
        hipGraph_t graph = nullptr;
        auto capture = [&](bool events) {
            h.graph_events = events;
            HIPC(hipStreamBeginCapture(h.stream, hipStreamCaptureModeThreadLocal));
            try {
                enqueue_fused(h, half, true);
            } catch (...) {
                hipGraph_t dropped = nullptr;
                hipStreamEndCapture(h.stream, &dropped);
                if (dropped) hipGraphDestroy(dropped);
                throw;
            }
            HIPC(hipStreamEndCapture(h.stream, &graph));
        };
        try {
            capture(true);
        } catch (const Failure&) {
            (void)hipGetLastError();
            capture(false);  // this runtime cannot record event nodes: time the whole replay
        }
        HIPC(hipGraphInstantiate(&h.graph_exec[half], graph, nullptr, nullptr, 0));
        HIPC(hipGraphDestroy(graph));
        h.graph_groups[half] = h.groups[half];
    }
    HIPC(hipGraphLaunch(h.graph_exec[half], h.stream));
}

void fill_timings(brp_hip& h, int half, uint32_t count) {
    auto& per = h.stage_ms_per_template[half];
    double* km = h.kernel_ms[half];
    per.assign(4 * size_t(count), 0.0);
    std::fill(km, km + 4, 0.0);
    float ms = 0;
    h.timed_launches[half] = h.timed_templates[half] = 0;
    if (h.opt.engine == BRP_ENGINE_PLAIN) {
        h.timed_launches[half] = h.timed_templates[half] = count;
        for (uint32_t r = 0; r < count; ++r)
            for (uint32_t s = 0; s < 4; ++s) {
                HIPC(hipEventElapsedTime(&ms, h.event(half, 5 * r + s), h.event(half, 5 * r + s + 1)));
                per[4 * r + s] = ms;
                km[s] += ms;
            }
        return;
    }
    if (!h.timed || (h.opt.graph && !h.opt.snapshot && !h.graph_events)) return;
    const auto& groups = h.groups[half];
    for (uint32_t gi = 0; gi < groups.size(); gi += h.stride) {
        const uint32_t first = groups[gi].first, n = groups[gi].count;
        ++h.timed_launches[half];
        h.timed_templates[half] += n;
        double stage[4] = {0, 0, 0, 0};
        HIPC(hipEventElapsedTime(&ms, h.event(half, 4 * gi), h.event(half, 4 * gi + 1)));
        stage[0] = ms;
        HIPC(hipEventElapsedTime(&ms, h.event(half, 4 * gi + 1), h.event(half, 4 * gi + 2)));
        stage[1] = ms;
        HIPC(hipEventElapsedTime(&ms, h.event(half, 4 * gi + 2), h.event(half, 4 * gi + 3)));
        stage[3] = ms;
        for (int s = 0; s < 4; ++s) km[s] += stage[s];
        for (uint32_t r = first; r < first + n; ++r)
            for (int s = 0; s < 4; ++s) per[4 * r + s] = stage[s] / n;
    }
}

// Enqueues the templates of `half` (parameters up, kernels, packed results down) and
// records its completion event; nothing waits.
void launch(brp_hip& h, int half) {
    (void)hipGetLastError();  // so that the check after the launches sees only theirs
    const uint32_t count = uint32_t(h.templates[half].size());
    const size_t base = h.slot_base(half);
    for (uint32_t i = 0; i < count; ++i) {
        h.params_host[base + i] = h.templates[half][i];
        h.params_host[base + i].result_slot = uint32_t(base + i);
    }
    const auto submit0 = std::chrono::steady_clock::now();
    // The previous retain of this half reads its slots and parameters: wait for it.
    if (h.device_top && h.retain_pending[half]) HIPC(hipStreamWaitEvent(h.stream, h.retained[half], 0));
    HIPC(hipMemcpyAsync(h.params_dev + base, h.params_host + base, sizeof(brp_params) * count,
                        hipMemcpyHostToDevice, h.stream));
    if (h.deferred[half]) {
        HIPC(hipMemsetAsync(h.steps_error_dev + half, 0, sizeof(uint32_t), h.stream));
        brp_engine::search_steps(h.stream, h.params_dev + base, count, h.lut, h.steps_error_dev + half);
    }
    if (h.device_top) {
        HIPC(hipMemcpyAsync(h.slot_table_dev + base, h.slot_table_host + base, sizeof(uint32_t) * count,
                            hipMemcpyHostToDevice, h.stream));
        HIPC(hipMemcpyAsync(h.slot_orbit_dev + base, h.slot_orbit_host + base, sizeof(float4) * count,
                            hipMemcpyHostToDevice, h.stream));
        brp_engine::toplist_thresholds(h.stream, h.params_dev + base, count, h.slot_table_dev + base,
                                       h.dtop_lowest[half]);
    }
    if (h.opt.engine == BRP_ENGINE_PLAIN) {
        require(half == 0, "the plain engine runs one batch at a time");
        if (h.opt.snapshot)
            HIPC(hipMemsetAsync(h.power_buf + h.base.n / 2, 0, sizeof(float) * 5ull * h.base.fundamental, h.stream));
        for (uint32_t r = 0; r < count; ++r) enqueue_plain(h, r);
        brp_engine::pack_results(h.stream, h.results, h.slot_words, count, h.keep, h.staging_dev);
    } else {
        if (h.opt.snapshot) {
            require(half == 0, "snapshots are kept for one batch at a time");
            HIPC(hipMemsetAsync(h.fa.snap_summed, 0, sizeof(float) * size_t(count) * 5 * h.fa.snap_f, h.stream));
            HIPC(hipMemsetAsync(h.fa.snap_X, 0, sizeof(float2) * size_t(count) * h.fa.snap_m1, h.stream));
        }
        run_fused(h, half);
    }
    // A launch the runtime rejected (configuration, resources) leaves its error here only.
    HIPC(hipGetLastError());
    if (h.device_top) {
        // Retains run in batch order on top_stream, each after its batch's kernels.
        HIPC(hipEventRecord(h.kernels_done[half], h.stream));
        HIPC(hipStreamWaitEvent(h.top_stream, h.kernels_done[half], 0));
        brp_engine::toplist_retain(h.top_stream, h.results + base * h.slot_words, h.slot_words, h.params_dev + base,
                                   count, h.slot_table_dev + base, h.slot_orbit_dev + base, h.dtop_tables, h.dtop,
                                   h.dtop_rows, h.dtop_lowest[half], h.dtop_error);
        HIPC(hipGetLastError());
        HIPC(hipEventRecord(h.retained[half], h.top_stream));
        h.retain_pending[half] = true;
    }
    h.submit_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - submit0).count();
    HIPC(hipMemcpyAsync(h.staging_host + base * (h.keep + 1) * 4, h.staging_dev + base * (h.keep + 1) * 4,
                        sizeof(uint32_t) * size_t(count) * (h.keep + 1) * 4, hipMemcpyDeviceToHost, h.stream));
    if (h.deferred[half]) {  // the lengths as run, and the length kernel's error word
        HIPC(hipMemcpyAsync(h.params_host + base, h.params_dev + base, sizeof(brp_params) * count,
                            hipMemcpyDeviceToHost, h.stream));
        HIPC(hipMemcpyAsync(h.steps_error_host + half, h.steps_error_dev + half, sizeof(uint32_t),
                            hipMemcpyDeviceToHost, h.stream));
    }
    if (!h.done[half]) HIPC(hipEventCreateWithFlags(&h.done[half], hipEventDisableTiming));
    HIPC(hipEventRecord(h.done[half], h.stream));
    h.ran[half] = false;
}

void finish(brp_hip& h, int half) {
    HIPC(hipEventSynchronize(h.done[half]));
    h.ran[half] = true;
    const uint32_t count = uint32_t(h.templates[half].size());
    if (h.deferred[half]) {
        require(!h.steps_error_host[half], "template produces empty series");
        const size_t base = h.slot_base(half);
        for (uint32_t k = 0; k < count; ++k) h.templates[half][k].steps = h.params_host[base + k].steps;
    }
    fill_timings(h, half, count);
    for (int s = 0; s < 4; ++s) h.total_ms[s] += h.kernel_ms[half][s];
    h.total_launches += h.timed_launches[half];
    h.total_templates += h.timed_templates[half];
}

void read_slot(brp_hip& h, int half, uint32_t slot, std::vector<brp_candidate>& out) {
    require(h.ran[half], "run() has not been called");
    require(slot < h.templates[half].size(), "candidate slot outside executed batch");
    const size_t global = h.slot_base(half) + slot;
    const uint32_t* rec = h.staging_host + global * (h.keep + 1) * 4;
    const uint32_t count = rec[0], flags = rec[1];
    const brp_params& p = h.templates[half][slot];
    if (flags || count > p.capacity)
        throw Failure(BRP_E_OVERFLOW, "GPU invalid resample index or candidate overflow; no valid result");
    out.resize(count);
    if (count <= h.keep) {
        std::memcpy(out.data(), rec + 4, sizeof(brp_candidate) * count);
    } else {
        HIPC(hipMemcpy(out.data(), h.results + global * h.slot_words + 4, sizeof(brp_candidate) * count,
                       hipMemcpyDeviceToHost));
    }
    std::sort(out.begin(), out.end(), [](const brp_candidate& a, const brp_candidate& b) {
        return a.level != b.level ? a.level < b.level : a.bin < b.bin;
    });
}

}  // namespace

extern "C" {

const char* brp_last_error(void) { return g_error.c_str(); }

int brp_hip_create(const brp_params* p, const float* input, size_t count, const brp_options* opt, brp_hip** out) {
    return guarded([&] {
        require(p && input && opt && out, "null argument");
        *out = nullptr;
        check_options(*p, *opt);
        require(count == p->unpadded, "input length mismatch");
        require(p->n >= 64 && p->steps && p->steps <= p->unpadded && p->unpadded <= p->n, "invalid template");
        int ndev = 0;
        HIPC(hipGetDeviceCount(&ndev));
        require(ndev > 0, "no HIP device");
        auto h = std::make_unique<brp_hip>();
        h->base = *p;
        h->opt = *opt;
        if (!h->opt.workgroup) h->opt.workgroup = 256;
        if (opt->engine == BRP_ENGINE_FUSED) h->base.flags |= 1u | 64u;  // packed, fused power
        HIPC(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        // Four zero guard samples after the observation: the fused gather reads three
        // consecutive samples starting at any index below unpadded.
        h->x = h->alloc<float>(size_t(p->unpadded) + 4);
        HIPC(hipMemcpy(h->x, input, sizeof(float) * p->unpadded, hipMemcpyHostToDevice));
        HIPC(hipMemset(h->x + p->unpadded, 0, 4 * sizeof(float)));
        std::vector<float2> lut(65);
        for (int i = 0; i < 65; ++i) lut[i] = make_float2(brp_lut::kTable[2 * i], brp_lut::kTable[2 * i + 1]);
        h->lut = h->upload(lut);
        h->slot_words = 4 + 4 * p->capacity;
        const size_t slots = 2 * size_t(opt->batch);  // two halves
        h->results = h->alloc<uint32_t>(slots * h->slot_words);
        HIPC(hipMemset(h->results, 0, sizeof(uint32_t) * slots * h->slot_words));
        h->staging_dev = h->alloc<uint32_t>(slots * (h->keep + 1) * 4);
        HIPC(hipHostMalloc(reinterpret_cast<void**>(&h->staging_host), sizeof(uint32_t) * slots * (h->keep + 1) * 4,
                           hipHostMallocDefault));
        h->params_dev = h->alloc<brp_params>(slots);
        HIPC(hipHostMalloc(reinterpret_cast<void**>(&h->params_host), sizeof(brp_params) * slots, hipHostMallocDefault));
        if (opt->engine == BRP_ENGINE_PLAIN) create_plain(*h);
        else create_fused(*h);
        h->templates[0].assign(1, *p);
        *out = h.release();
    });
}

int brp_hip_destroy(brp_hip* h) {
    return guarded([&] { delete h; });
}

int brp_hip_set_templates(brp_hip* h, const brp_params* p, size_t count) {
    return guarded([&] {
        require(h && p, "null argument");
        require(count && count <= h->opt.batch, "template batch exceeds retained result capacity");
        for (size_t i = 0; i < count; ++i) check_template(*h, p[i]);
        h->templates[0].assign(p, p + count);
        h->ran[0] = false;
    });
}

int brp_hip_run(brp_hip* h, brp_timings* t, double* template_stage_ms) {
    return guarded([&] {
        require(h, "null handle");
        const auto start = std::chrono::steady_clock::now();
        const double cpu0 = process_cpu_ms();
        launch(*h, 0);
        finish(*h, 0);
        if (t) {
            t->submit_ms = h->submit_ms;
            for (int s = 0; s < 4; ++s) t->stage_ms[s] = h->kernel_ms[0][s];
            t->cpu_ms = process_cpu_ms() - cpu0;
            t->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
        }
        if (template_stage_ms)
            std::copy(h->stage_ms_per_template[0].begin(), h->stage_ms_per_template[0].end(), template_stage_ms);
    });
}

int brp_hip_candidates(brp_hip* h, uint32_t slot, brp_candidate* out, size_t cap, size_t* count) {
    return guarded([&] {
        require(h && count, "null argument");
        std::vector<brp_candidate> v;
        read_slot(*h, 0, slot, v);
        *count = v.size();
        require(v.size() <= cap || !out, "candidate buffer too small");
        if (out) std::copy(v.begin(), v.end(), out);
    });
}

int brp_hip_snapshot(brp_hip* h, uint32_t slot, float* resampled, int32_t* index, float* fft, float* power,
                     float* summed) {
    return guarded([&] {
        require(h && h->ran[0], "run() has not been called");
        require(slot < h->templates[0].size(), "snapshot slot outside executed batch");
        const brp_params& p = h->templates[0][slot];
        const uint32_t n = p.n, half = n / 2;
        if (h->opt.engine == BRP_ENGINE_PLAIN) {
            // The plain engine runs a batch sequentially on shared buffers (like the reference),
            // so its arrays hold the last template of the batch.
            require(slot + 1 == h->templates[0].size(), "plain engine snapshots hold the last template of the batch");
            if (resampled) {
                HIPC(hipMemcpy(resampled, h->pb.raw, sizeof(float) * p.steps, hipMemcpyDeviceToHost));
                float mean = 0;
                HIPC(hipMemcpy(&mean, h->pb.mean, sizeof(float), hipMemcpyDeviceToHost));
                std::fill(resampled + p.steps, resampled + n, mean);
            }
            if (index) {
                require(h->pb.index, "index snapshot needs options.snapshot");
                HIPC(hipMemcpy(index, h->pb.index, sizeof(int32_t) * p.steps, hipMemcpyDeviceToHost));
            }
            if (fft) HIPC(hipMemcpy(fft, h->pb.X, sizeof(float2) * (half + 1), hipMemcpyDeviceToHost));
            if (power) HIPC(hipMemcpy(power, h->power_buf, sizeof(float) * half, hipMemcpyDeviceToHost));
            if (summed) HIPC(hipMemcpy(summed, h->power_buf + half, sizeof(float) * 5ull * p.fundamental, hipMemcpyDeviceToHost));
            return;
        }
        require(h->opt.snapshot, "fused engine snapshots need options.snapshot");
        const auto& a = h->fa;
        if (resampled) {
            HIPC(hipMemcpy(resampled, a.snap_resampled + size_t(slot) * a.snap_n, sizeof(float) * p.steps,
                           hipMemcpyDeviceToHost));
            float mean = 0;
            HIPC(hipMemcpy(&mean, a.mean + slot, sizeof(float), hipMemcpyDeviceToHost));
            std::fill(resampled + p.steps, resampled + n, mean);
        }
        if (index)
            HIPC(hipMemcpy(index, a.snap_index + size_t(slot) * a.snap_n, sizeof(int32_t) * p.steps, hipMemcpyDeviceToHost));
        if (fft)
            HIPC(hipMemcpy(fft, a.snap_X + size_t(slot) * a.snap_m1, sizeof(float2) * (half + 1), hipMemcpyDeviceToHost));
        if (power) {
            // The slot's launch group keeps its power at first * M_max + (slot - first) * M.
            const brp_hip::Group* grp = nullptr;
            for (const auto& g : h->groups[0])
                if (slot >= g.first && slot < g.first + g.count) grp = &g;
            require(grp != nullptr, "snapshot slot outside executed batch");
            const size_t at = size_t(grp->first) * h->m_max + size_t(slot - grp->first) * h->series[grp->series].shape.m;
            HIPC(hipMemcpy(power, a.P + at, sizeof(float) * half, hipMemcpyDeviceToHost));
        }
        if (summed)  // level h of the slot at h * snap_f (the largest F of the handle)
            for (size_t lv = 0; lv < 5; ++lv)
                HIPC(hipMemcpy(summed + lv * p.fundamental, a.snap_summed + (size_t(slot) * 5 + lv) * a.snap_f,
                               sizeof(float) * p.fundamental, hipMemcpyDeviceToHost));
    });
}

size_t brp_hip_device_info(brp_hip* h, char* buf, size_t cap) {
    std::ostringstream s;
    int dev = 0;
    hipDeviceProp_t prop{};
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        int rt = 0;
        hipRuntimeGetVersion(&rt);
        s << prop.name << "; " << prop.gcnArchName << "; CUs=" << prop.multiProcessorCount
          << "; wavefront=" << prop.warpSize << "; LDS/block=" << prop.sharedMemPerBlock
          << "; HBM=" << (prop.totalGlobalMem >> 20) << " MiB; clock_khz=" << prop.clockRate << "; HIP runtime " << rt;
    } else {
        s << "no HIP device";
    }
    if (h && h->opt.engine == BRP_ENGINE_FUSED) {
        s << "; engine=fused";
        for (const auto& ser : h->series)
            s << " N=" << ser.n << " C=" << ser.shape.c << " R=" << ser.shape.r << " S1=" << ser.shape.s1
              << " S2=" << ser.shape.s2 << " T1=" << ser.shape.t1 << " T2=" << ser.shape.t2
              << " exchange=" << (ser.shape.mirror ? "mirror" : "plain") << ";";
        s << " group=" << h->group << " lanes=" << h->lanes;
    }
    else if (h)
        s << "; engine=plain N=" << h->base.n;
    const std::string text = s.str();
    if (buf && cap) {
        const size_t n = std::min(cap - 1, text.size());
        std::memcpy(buf, text.data(), n);
        buf[n] = 0;
    }
    return text.size();
}

uint64_t brp_hip_allocated_bytes(const brp_hip* h) { return h ? h->allocated : 0; }

}  // extern "C"

namespace {
// ------------------------------------------------------------- counters --
std::string json_quote(const std::string& text) {
    std::string out = "\"";
    for (char c : text) {
        if (c == '"' || c == '\\') out += '\\';
        if (c == '\n') out += "\\n";
        else if (uint8_t(c) >= 0x20) out += c;
    }
    return out + "\"";
}

// The kernel's demangled name without its namespace and parameter list.
std::string kernel_label(const std::string& symbol) {
    int status = 0;
    char* d = abi::__cxa_demangle(symbol.c_str(), nullptr, nullptr, &status);
    std::string name = (status == 0 && d) ? d : symbol;
    std::free(d);
    // Function templates demangle with their return type first ("void (anonymous
    // namespace)::k_pass1<...>(...)").
    if (name.compare(0, 5, "void ") == 0) name = name.substr(5);
    for (const char* ns : {"(anonymous namespace)::", "brp_engine::"})
        for (size_t at; (at = name.find(ns)) != std::string::npos;) name.erase(at, std::strlen(ns));
    int depth = 0;
    for (size_t i = 0; i < name.size(); ++i) {  // cut the parameter list at depth 0
        if (name[i] == '<') ++depth;
        else if (name[i] == '>') --depth;
        else if (name[i] == '(' && depth == 0) return name.substr(0, i);
    }
    return name;
}

// The reference's stage of a kernel (src/vulkan.cpp:221-266; the fused kernels span stages).
std::string kernel_stage(const std::string& label) {
    auto starts = [&](const char* p) { return label.compare(0, std::strlen(p), p) == 0; };
    if (starts("k_pass1")) return "pass1_resample_prepare_fft";
    if (starts("k_pass2")) return "pass2_fft_power";
    if (starts("k_harmonic") || starts("k_hs_")) return "harmonic_candidates";
    if (starts("k_resample") || starts("k_mean") || starts("k_pack<") || label == "k_pack") return "resample_prepare";
    if (starts("k_stockham")) return "fft";
    if (starts("k_real") || starts("k_power")) return "power";
    if (starts("brp_copy")) return "calibration_copy";
    return "other";
}

// profile_counters (src/vulkan.cpp:346-406) in process: the handle's last batch replayed once
// per counter set under rocprofiler-sdk's dispatch counting service, with a plain copy of known
// size in the same replay that calibrates FETCH_SIZE / WRITE_SIZE to bytes.  Each set fits one
// hardware pass on gfx950 (at most 8 SQ, 4 TCC, 4 TCP, 2 TA, 2 TD, 2 GRBM counters).
std::string profile_counters_json(brp_hip* h) {
    const std::string why = brp_counters::unavailable_reason();
    if (!why.empty()) return "{\"available\":false,\"reason\":" + json_quote(why) + "}";
    require(h, "null handle");
    require(h->ran[0], "run() has not been called");
    const std::vector<std::vector<std::string>> sets = {
        {"SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR", "SQ_INSTS_LDS", "SQ_WAVE_CYCLES",
         "SQ_WAIT_INST_ANY", "GRBM_GUI_ACTIVE", "FETCH_SIZE"},
        {"WRITE_SIZE", "TA_TA_BUSY_sum", "TD_TD_BUSY_sum", "TCP_TOTAL_CACHE_ACCESSES_sum", "TCC_HIT_sum", "TCC_MISS_sum"}};
    const size_t copy_bytes = size_t(64) << 20;
    void* src = nullptr;
    void* dst = nullptr;
    HIPC(hipMalloc(&src, copy_bytes));
    HIPC(hipMalloc(&dst, copy_bytes));
    struct Agg {
        std::string stage;
        uint64_t dispatches = 0;
        double us = 0;
        std::map<std::string, double> sum;
    };
    std::map<std::string, Agg> kernels;
    std::vector<std::string> missing, gone;
    try {
        for (const auto& set : sets) {
            const auto ds = brp_counters::collect(
                set,
                [&] {
                    launch(*h, 0);
                    finish(*h, 0);
                    brp_engine::copy16(h->stream, src, dst, copy_bytes);
                    HIPC(hipStreamSynchronize(h->stream));
                },
                &gone);
            missing.insert(missing.end(), gone.begin(), gone.end());
            std::map<std::string, uint64_t> seen;
            for (const auto& d : ds) {
                const std::string label = kernel_label(d.kernel);
                Agg& a = kernels[label];
                a.stage = kernel_stage(label);
                if (&set == &sets[0]) {
                    ++a.dispatches;
                    a.us += double(d.end_ns - d.start_ns) * 1e-3;
                }
                for (const auto& kv : d.values) a.sum[kv.first] += kv.second;
            }
        }
    } catch (...) {
        hipFree(src);
        hipFree(dst);
        throw;
    }
    HIPC(hipFree(src));
    HIPC(hipFree(dst));
    // Counter KiB -> bytes, calibrated on the copy (which reads and writes copy_bytes).
    double fetch_factor = 0, write_factor = 0;
    for (const auto& kv : kernels) {
        if (kv.second.stage != "calibration_copy" || !kv.second.dispatches) continue;
        const auto& m = kv.second.sum;
        const double n = double(kv.second.dispatches);
        if (m.count("FETCH_SIZE") && m.at("FETCH_SIZE") > 0) fetch_factor = double(copy_bytes) / (m.at("FETCH_SIZE") / n);
        if (m.count("WRITE_SIZE") && m.at("WRITE_SIZE") > 0) write_factor = double(copy_bytes) / (m.at("WRITE_SIZE") / n);
    }
    std::ostringstream out;
    out.precision(10);
    out << "{\"available\":true,\"source\":\"rocprofiler-sdk dispatch counting service\",\"passes\":" << sets.size()
        << ",\"templates\":" << h->templates[0].size() << ",\"calibration\":{\"copy_bytes\":" << copy_bytes
        << ",\"fetch_bytes_per_unit\":" << fetch_factor << ",\"write_bytes_per_unit\":" << write_factor << "},\"missing\":[";
    for (size_t i = 0; i < missing.size(); ++i) out << (i ? "," : "") << json_quote(missing[i]);
    out << "],\"stages\":[";
    bool first = true;
    for (const auto& kv : kernels) {
        const Agg& a = kv.second;
        if (!a.dispatches) continue;
        const double n = double(a.dispatches);
        out << (first ? "" : ",") << "{\"stage\":" << json_quote(a.stage) << ",\"kernel\":" << json_quote(kv.first)
            << ",\"dispatches\":" << a.dispatches << ",\"counters\":{";
        first = false;
        bool f2 = true;
        for (const auto& c : a.sum) {
            out << (f2 ? "" : ",") << json_quote(c.first) << ":" << c.second / n;
            f2 = false;
        }
        out << "},\"derived\":{\"duration_us\":" << a.us / n;
        auto get = [&](const char* name) { return a.sum.count(name) ? a.sum.at(name) / n : -1.0; };
        if (get("FETCH_SIZE") >= 0 && fetch_factor > 0) out << ",\"read_bytes\":" << get("FETCH_SIZE") * fetch_factor;
        if (get("WRITE_SIZE") >= 0 && write_factor > 0) out << ",\"write_bytes\":" << get("WRITE_SIZE") * write_factor;
        if (get("SQ_WAVES") > 0 && get("SQ_INSTS_VALU") >= 0)
            out << ",\"valu_per_wave\":" << get("SQ_INSTS_VALU") / get("SQ_WAVES");
        out << "}}";
    }
    out << "]}";
    return out.str();
}

}  // namespace

extern "C" {

size_t brp_hip_profile_counters(brp_hip* h, char* buf, size_t cap) {
    std::string text;
    try {
        text = profile_counters_json(h);
    } catch (const std::exception& e) {
        text = "{\"available\":false,\"reason\":" + json_quote(e.what()) + "}";
    }
    if (buf && cap) {
        const size_t n = std::min(cap - 1, text.size());
        std::memcpy(buf, text.data(), n);
        buf[n] = 0;
    }
    return text.size();
}

int brp_hip_kernel_ms(brp_hip* h, double out[4], uint64_t* launches, uint64_t* templates) {
    return guarded([&] {
        require(h && out, "null argument");
        for (int s = 0; s < 4; ++s) { out[s] = h->total_ms[s]; h->total_ms[s] = 0; }
        if (launches) *launches = h->total_launches;
        if (templates) *templates = h->total_templates;
        h->total_launches = h->total_templates = 0;
    });
}

int brp_hip_autotune(const brp_params* templates, size_t count, const float* input, size_t input_count,
                     const brp_options* base, uint32_t reps, brp_options* best, char* report, size_t cap) {
    return guarded([&] {
        require(templates && count && input && base && best, "null argument");
        require(reps > 0 && reps <= 1000, "invalid repetitions");
        const uint32_t n = templates[0].n;
        // The compiled shapes of the length; with the precision variant, those that have the
        // radix-16 stages it runs on the matrix cores.
        std::vector<brp_engine::FusedShape> shapes;
        for (const auto& s : brp_engine::fused_shapes(n))
            if (!base->precision || brp_engine::fused_mfma_capable(s, base->precision)) shapes.push_back(s);
        require(!shapes.empty(), base->precision ? "split-FP16 variant: no shape of this length has the radix-16 "
                                                   "stages it asks for"
                                                 : "fused engine: unsupported series length");
        std::ostringstream log;
        double best_us = 1e300;
        *best = *base;
        auto configure = [&](const brp_engine::FusedShape& s, uint32_t group, uint32_t lanes) {
            brp_options o = *base;
            o.engine = BRP_ENGINE_FUSED;
            o.packed = 1;
            o.snapshot = 0;
            o.batch = uint32_t(count);
            o.split_c = s.c;
            o.split_s1 = s.s1;
            o.split_s2 = s.s2;
            o.split_t1 = s.t1;
            o.split_t2 = s.t2;
            o.group = group;
            o.lanes = lanes;
            brp_hip* h = nullptr;
            if (int rc = brp_hip_create(&templates[0], input, input_count, &o, &h)) throw Failure(rc, brp_last_error());
            std::unique_ptr<brp_hip> owner(h);
            if (int rc = brp_hip_set_templates(h, templates, count)) throw Failure(rc, brp_last_error());
            return owner;
        };
        {
            // Untimed runs of the first shape, so that the device has left its idle clocks
            // before the first configuration is timed.
            auto warm = configure(shapes[0], std::min<uint32_t>(8u, uint32_t(count)), 1u);
            for (int r = 0; r < 3; ++r) {
                launch(*warm, 0);
                finish(*warm, 0);
            }
        }
        for (const auto& s : shapes) {
            for (uint32_t group : {2u, 4u, 8u, 16u})
                for (uint32_t lanes : {1u, 2u}) {
                    if (group > count || (lanes == 2 && 2 * group > count)) continue;
                    if (base->lanes && lanes != base->lanes) continue;  // lanes fixed by the caller
                    std::unique_ptr<brp_hip> owner = configure(s, group, lanes);
                    brp_hip* h = owner.get();
                    launch(*h, 0);
                    finish(*h, 0);  // warm-up (and graph capture)
                    std::vector<double> us;
                    for (uint32_t r = 0; r < reps; ++r) {
                        const auto t0 = std::chrono::steady_clock::now();
                        launch(*h, 0);
                        finish(*h, 0);
                        us.push_back(std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0)
                                         .count() / double(count));
                    }
                    std::sort(us.begin(), us.end());
                    const double med = us[us.size() / 2];
                    log << "{\"C\":" << s.c << ",\"R\":" << s.r << ",\"S1\":" << s.s1 << ",\"S2\":" << s.s2
                        << ",\"T1\":" << s.t1 << ",\"T2\":" << s.t2
                        << ",\"group\":" << group << ",\"lanes\":" << lanes << ",\"us_per_template\":" << med << "}\n";
                    if (med < best_us) {
                        best_us = med;
                        best->split_c = s.c;
                        best->split_s1 = s.s1;
                        best->split_s2 = s.s2;
                        best->split_t1 = s.t1;
                        best->split_t2 = s.t2;
                        best->group = group;
                        best->lanes = lanes;
                    }
                }
        }
        const std::string text = log.str();
        if (report && cap) {
            const size_t m = std::min(cap - 1, text.size());
            std::memcpy(report, text.data(), m);
            report[m] = 0;
        }
    });
}

int brp_hip_copy_bandwidth(size_t bytes, int reps, double* gbps) {
    return guarded([&] {
        require(gbps && bytes >= 16 && reps > 0, "invalid copy calibration");
        (void)hipGetLastError();
        bytes &= ~size_t(15);
        void *a = nullptr, *b = nullptr;
        hipStream_t st = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto release = [&] {
            if (st) hipStreamSynchronize(st);
            if (e0) hipEventDestroy(e0);
            if (e1) hipEventDestroy(e1);
            if (st) hipStreamDestroy(st);
            if (a) hipFree(a);
            if (b) hipFree(b);
        };
        try {
            HIPC(hipMalloc(&a, bytes));
            HIPC(hipMalloc(&b, bytes));
            HIPC(hipMemset(a, 1, bytes));
            HIPC(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            HIPC(hipEventCreate(&e0));
            HIPC(hipEventCreate(&e1));
            brp_engine::copy16(st, a, b, bytes);  // warm-up
            HIPC(hipEventRecord(e0, st));
            for (int r = 0; r < reps; ++r) brp_engine::copy16(st, a, b, bytes);
            HIPC(hipGetLastError());
            HIPC(hipEventRecord(e1, st));
            HIPC(hipEventSynchronize(e1));
            float ms = 0;
            HIPC(hipEventElapsedTime(&ms, e0, e1));
            *gbps = 2.0 * double(bytes) * reps / (double(ms) * 1e-3) / 1e9;
        } catch (...) {
            release();
            throw;
        }
        release();
    });
}

int brp_hip_preprocess(const uint8_t* bytes, size_t samples, uint32_t bits, double scale, uint32_t n,
                       uint32_t window, float rfi_ratio, float* out, float* decoded, uint32_t* rfi_bins) {
    return brp_hip_preprocess_zap(bytes, samples, bits, scale, n, window, rfi_ratio, nullptr, 0, 0.0, 1.0, out,
                                  decoded, rfi_bins);
}

int brp_hip_preprocess_zap(const uint8_t* bytes, size_t samples, uint32_t bits, double scale, uint32_t n,
                           uint32_t window, float rfi_ratio, const double* zap, size_t zap_ranges, double sample_us,
                           double padding, float* out, float* decoded, uint32_t* rfi_bins) {
    return guarded([&] {
        require(bytes && out, "null argument");
        require(bits == 4 || bits == 8, "unsupported DD sample format");
        require(!(bits == 4 && (samples & 1)), "odd packed DD length");
        require(samples && samples <= n && std::isfinite(scale) && scale > 0, "invalid DD metadata");
        require(n >= 64 && n <= 3u * (1u << 22) && n % 2 == 0 && (pow2(n) || (n % 3 == 0 && pow2(n / 3))),
                "observation exceeds supported length");
        require(window && !(window & 1) && window <= n / 2 && window <= 3072, "unsupported whitening window");
        require(std::isfinite(rfi_ratio) && rfi_ratio >= 0, "invalid RFI ratio");
        require(!zap_ranges || (zap && std::isfinite(sample_us) && sample_us > 0 && std::isfinite(padding) &&
                                padding >= 1 && padding <= 10),
                "invalid zap arguments");
        int ndev = 0;
        HIPC(hipGetDeviceCount(&ndev));
        require(ndev > 0, "no HIP device");
        brp_engine::preprocess_observation(bytes, samples, bits, scale, n, window, rfi_ratio, zap, zap_ranges,
                                           sample_us, padding, out, decoded, rfi_bins);
    });
}

int brp_hip_add_series(brp_hip* h, const brp_params* p, const float* input, size_t count) {
    return guarded([&] {
        require(h && p && input, "null argument");
        require(h->opt.engine == BRP_ENGINE_FUSED, "more than one series needs the fused engine");
        require(count == p->unpadded, "input length mismatch");
        require(p->n >= 64 && p->steps && p->steps <= p->unpadded && p->unpadded <= p->n, "invalid template");
        HIPC(hipStreamSynchronize(h->stream));
        float* x = h->alloc<float>(size_t(p->unpadded) + 4);
        HIPC(hipMemcpy(x, input, sizeof(float) * p->unpadded, hipMemcpyHostToDevice));
        HIPC(hipMemset(x + p->unpadded, 0, 4 * sizeof(float)));
        add_fused_series(*h, *p, x);
    });
}

}  // extern "C"

namespace {

// The bank loop of src/workunit_main.cpp:77-136 over prepared templates, each retained
// into the toplist of its series (tops[series]).  Series are independent, so the bank
// runs series by series (template order kept within each): batches of options.batch
// templates are then whole launch groups of one length, except where two series meet,
// and the launch plans repeat, so captured graphs are replayed.  Thresholds are frozen
// at each batch start and retain() replays template order within each series.
void bank_loop(brp_hip& h, const std::vector<brp_params>& base, const brp_orbit* orbits, brp_toplist* const* tops,
               size_t ntops, uint32_t* cand_count, brp_params* templates_out, double* gpu_ms,
               bool steps_deferred = false) {
    const size_t count = base.size();
    std::vector<uint32_t> ser(count, 0);
    for (size_t i = 0; i < count; ++i) {
        check_template(h, base[i], steps_deferred);
        if (h.opt.engine == BRP_ENGINE_FUSED) ser[i] = series_of(h, base[i]);
        require(ser[i] < ntops && tops[ser[i]], "no toplist for a template's series");
    }
    std::vector<size_t> seq(count);
    for (size_t i = 0; i < count; ++i) seq[i] = i;
    std::stable_sort(seq.begin(), seq.end(), [&](size_t x, size_t y) { return ser[x] < ser[y]; });
    const size_t batch = h.opt.batch;
    const size_t batches = (count + batch - 1) / batch;
    // The fused engine keeps two batches in flight: batch b + 1 is enqueued with the
    // thresholds of the toplists before batch b is retained.  Those are never higher
    // than the thresholds batch b + 1 would see one batch at a time, so its candidates
    // are a superset, and retain() with each template's current thresholds keeps
    // exactly the same rows.  The plain engine runs batch by batch, as the reference.
    const bool overlap = h.opt.engine == BRP_ENGINE_FUSED && !h.opt.snapshot;
    auto enqueue = [&](size_t b) {
        const int half = overlap ? int(b & 1) : 0;
        const size_t first = b * batch, n = std::min(batch, count - first);
        h.templates[half].resize(n);
        h.deferred[half] = false;
        for (size_t k = 0; k < n; ++k) {
            const size_t i = seq[first + k];
            h.templates[half][k] = base[i];
            h.deferred[half] = h.deferred[half] || !base[i].steps;
            brp_toplist_thresholds(tops[ser[i]], &h.templates[half][k]);
        }
        launch(h, half);
    };
    double total_gpu = 0;
    std::vector<brp_candidate> cand;
    if (batches) enqueue(0);
    for (size_t b = 0; b < batches; ++b) {
        const int half = overlap ? int(b & 1) : 0;
        if (overlap && b + 1 < batches) enqueue(b + 1);
        finish(h, half);
        for (int s = 0; s < 4; ++s) total_gpu += h.kernel_ms[half][s];
        const size_t first = b * batch, n = h.templates[half].size();
        for (size_t k = 0; k < n; ++k) {
            const size_t i = seq[first + k];
            brp_params now = base[i];
            now.steps = h.templates[half][k].steps;  // as run (computed on the device when deferred)
            brp_toplist* top = tops[ser[i]];
            brp_toplist_thresholds(top, &now);  // the toplist as of this template
            read_slot(h, half, uint32_t(k), cand);
            if (int rc = brp_toplist_retain(top, cand.data(), cand.size(), &orbits[i], &now))
                throw Failure(rc, brp_last_error());
            if (templates_out) templates_out[i] = now;
            if (cand_count) cand_count[i] = uint32_t(cand.size());
        }
        if (!overlap && b + 1 < batches) enqueue(b + 1);
    }
    h.deferred[0] = h.deferred[1] = false;
    if (gpu_ms) *gpu_ms = total_gpu;
}

// Device toplist state for `tables` toplists of `rows` rows (allocated once, grown as needed).
void ensure_device_toplist(brp_hip& h, uint32_t tables, uint32_t rows) {
    const size_t slots = 2 * size_t(h.opt.batch);
    if (!h.slot_table_dev) {
        h.slot_table_dev = h.alloc<uint32_t>(slots);
        h.slot_orbit_dev = h.alloc<float4>(slots);
        h.dtop_error = h.alloc<uint32_t>(1);
        HIPC(hipHostMalloc(reinterpret_cast<void**>(&h.slot_table_host), sizeof(uint32_t) * slots, hipHostMallocDefault));
        HIPC(hipHostMalloc(reinterpret_cast<void**>(&h.slot_orbit_host), sizeof(float4) * slots, hipHostMallocDefault));
        HIPC(hipStreamCreateWithFlags(&h.top_stream, hipStreamNonBlocking));
        for (int k = 0; k < 2; ++k) {
            HIPC(hipEventCreateWithFlags(&h.kernels_done[k], hipEventDisableTiming));
            HIPC(hipEventCreateWithFlags(&h.retained[k], hipEventDisableTiming));
        }
    }
    const uint32_t need = tables * rows;
    if (need > h.dtop_capacity) {
        h.release(h.dtop);
        h.release(h.dtop_lowest[0]);
        h.release(h.dtop_lowest[1]);
        h.dtop = h.alloc<brp_engine::TopRow>(5 * size_t(need));
        h.dtop_lowest[0] = h.alloc<float>(5 * size_t(tables));
        h.dtop_lowest[1] = h.alloc<float>(5 * size_t(tables));
        h.dtop_capacity = need;
    }
    h.dtop_tables = tables;
    h.dtop_rows = rows;
    h.retain_pending[0] = h.retain_pending[1] = false;
}

void upload_toplists(brp_hip& h, brp_toplist* const* tops, size_t ntops) {
    std::vector<brp_engine::TopRow> rows(5 * size_t(h.dtop_rows) * ntops);
    std::vector<float> lowest(5 * ntops);
    for (size_t k = 0; k < ntops; ++k) {
        for (size_t r = 0; r < 5 * size_t(h.dtop_rows); ++r) {
            const brp_bank_row& b = tops[k]->table[r];
            rows[k * 5 * h.dtop_rows + r] =
                brp_engine::TopRow{float(b.power), float(b.period), float(b.tau), float(b.psi), b.bin, b.harmonics};
        }
        for (size_t lv = 0; lv < 5; ++lv) lowest[k * 5 + lv] = float(tops[k]->table[(lv + 1) * h.dtop_rows - 1].power);
    }
    HIPC(hipStreamSynchronize(h.top_stream));
    HIPC(hipMemcpyAsync(h.dtop, rows.data(), sizeof(brp_engine::TopRow) * rows.size(), hipMemcpyHostToDevice, h.stream));
    for (int half = 0; half < 2; ++half)
        HIPC(hipMemcpyAsync(h.dtop_lowest[half], lowest.data(), sizeof(float) * lowest.size(), hipMemcpyHostToDevice,
                            h.stream));
    HIPC(hipMemsetAsync(h.dtop_error, 0, sizeof(uint32_t), h.stream));
    HIPC(hipStreamSynchronize(h.stream));
}

void download_toplists(brp_hip& h, brp_toplist* const* tops, size_t ntops) {
    std::vector<brp_engine::TopRow> rows(5 * size_t(h.dtop_rows) * ntops);
    uint32_t error = 0;
    HIPC(hipStreamSynchronize(h.top_stream));
    HIPC(hipMemcpyAsync(rows.data(), h.dtop, sizeof(brp_engine::TopRow) * rows.size(), hipMemcpyDeviceToHost, h.stream));
    HIPC(hipMemcpyAsync(&error, h.dtop_error, sizeof(uint32_t), hipMemcpyDeviceToHost, h.stream));
    HIPC(hipStreamSynchronize(h.stream));
    if (error & 1u) throw Failure(BRP_E_OVERFLOW, "GPU invalid resample index or candidate overflow; no valid result");
    if (error & 2u) throw Failure(BRP_E_INVALID, "invalid compact candidate");
    if (error & 4u) throw Failure(BRP_E_RUNTIME, "device toplist: more than 4096 candidates of one level in one template");
    for (size_t k = 0; k < ntops; ++k)
        for (size_t r = 0; r < 5 * size_t(h.dtop_rows); ++r) {
            const brp_engine::TopRow& d = rows[k * 5 * h.dtop_rows + r];
            tops[k]->table[r] = brp_bank_row{double(d.power), double(d.period), double(d.tau), double(d.psi), 0.0,
                                             d.harmonics, d.bin};
        }
}

// The bank loop of bank_loop with the toplists on the device: the same launch sequence
// and two batches in flight, with thresholds raised and retain() run in the launch stream,
// so the host only collects candidate counts.  The toplists equal bank_loop's.
void bank_loop_device(brp_hip& h, const std::vector<brp_params>& base, const brp_orbit* orbits,
                      brp_toplist* const* tops, size_t ntops, uint32_t* cand_count, double* gpu_ms) {
    require(h.opt.engine == BRP_ENGINE_FUSED && !h.opt.snapshot, "device toplist: fused engine without snapshots");
    require(ntops > 0, "no toplist");
    const uint32_t rows = tops[0] ? tops[0]->rows : 0;
    for (size_t k = 0; k < ntops; ++k)
        require(tops[k] && tops[k]->rows == rows, "device toplist: every toplist needs the same number of rows");
    require(rows >= 1 && rows <= brp_engine::kDeviceToplistMaxRows, "device toplist: at most 2048 rows per level");
    const size_t count = base.size();
    std::vector<uint32_t> ser(count, 0);
    for (size_t i = 0; i < count; ++i) {
        check_template(h, base[i]);
        ser[i] = series_of(h, base[i]);
        require(ser[i] < ntops, "no toplist for a template's series");
    }
    std::vector<size_t> seq(count);
    for (size_t i = 0; i < count; ++i) seq[i] = i;
    std::stable_sort(seq.begin(), seq.end(), [&](size_t x, size_t y) { return ser[x] < ser[y]; });
    ensure_device_toplist(h, uint32_t(ntops), rows);
    upload_toplists(h, tops, ntops);
    struct Off {
        brp_hip& h;
        ~Off() {
            if (h.top_stream) hipStreamSynchronize(h.top_stream);
            h.device_top = false;
            h.retain_pending[0] = h.retain_pending[1] = false;
        }
    } off{h};
    h.device_top = true;
    const size_t batch = h.opt.batch;
    const size_t batches = (count + batch - 1) / batch;
    auto enqueue = [&](size_t b) {
        const int half = int(b & 1);
        const size_t first = b * batch, n = std::min(batch, count - first), at = h.slot_base(half);
        h.templates[half].resize(n);
        for (size_t k = 0; k < n; ++k) {
            const size_t i = seq[first + k];
            h.templates[half][k] = base[i];
            h.slot_table_host[at + k] = ser[i];
            h.slot_orbit_host[at + k] =
                make_float4(float(orbits[i].period), float(orbits[i].tau), float(orbits[i].psi), 0.0f);
        }
        launch(h, half);
    };
    double total_gpu = 0;
    if (batches) enqueue(0);
    for (size_t b = 0; b < batches; ++b) {
        const int half = int(b & 1);
        if (b + 1 < batches) enqueue(b + 1);
        finish(h, half);
        for (int s = 0; s < 4; ++s) total_gpu += h.kernel_ms[half][s];
        const size_t first = b * batch, n = h.templates[half].size();
        for (size_t k = 0; k < n; ++k) {
            const uint32_t* rec = h.staging_host + (h.slot_base(half) + k) * (h.keep + 1) * 4;
            if (rec[1] || rec[0] > h.templates[half][k].capacity)
                throw Failure(BRP_E_OVERFLOW, "GPU invalid resample index or candidate overflow; no valid result");
            if (cand_count) cand_count[seq[first + k]] = rec[0];
        }
    }
    download_toplists(h, tops, ntops);
    if (gpu_ms) *gpu_ms = total_gpu;
}

}  // namespace

extern "C" {

int brp_hip_bank_search_device(brp_hip* h, const brp_params* templates, const brp_orbit* orbits, size_t count,
                               brp_toplist* const* tops, size_t ntops, uint32_t* cand_count, double* gpu_ms) {
    return guarded([&] {
        require(h && templates && orbits && tops, "null argument");
        std::vector<brp_params> base(templates, templates + count);
        bank_loop_device(*h, base, orbits, tops, ntops, cand_count, gpu_ms);
    });
}

int brp_hip_toplist_retain_device(const brp_params* templates, const brp_orbit* orbits, size_t count,
                                  const brp_candidate* candidates, const uint32_t* cand_count, brp_toplist* top) {
    return guarded([&] {
        require(templates && orbits && top && (count == 0 || cand_count), "null argument");
        require(top->rows >= 1 && top->rows <= brp_engine::kDeviceToplistMaxRows,
                "device toplist: at most 2048 rows per level");
        int ndev = 0;
        HIPC(hipGetDeviceCount(&ndev));
        require(ndev > 0, "no HIP device");
        // Result slots in the engine's layout: {count, flags, 0, 0} then the candidates.
        uint32_t widest = 0;
        size_t total = 0;
        for (size_t i = 0; i < count; ++i) {
            widest = std::max(widest, cand_count[i]);
            total += cand_count[i];
        }
        require(!total || candidates, "null argument");
        const uint32_t words = 4 + 4 * widest;
        std::vector<uint32_t> slots(size_t(count) * words, 0u);
        std::vector<brp_params> params(templates, templates + count);
        std::vector<float4> orbit(count);
        size_t next = 0;
        for (size_t i = 0; i < count; ++i) {
            uint32_t* rec = slots.data() + i * words;
            rec[0] = cand_count[i];
            params[i].capacity = std::max(params[i].capacity, cand_count[i]);
            std::memcpy(rec + 4, candidates + next, sizeof(brp_candidate) * cand_count[i]);
            next += cand_count[i];
            orbit[i] = make_float4(float(orbits[i].period), float(orbits[i].tau), float(orbits[i].psi), 0.0f);
        }
        std::vector<uint32_t> table(count, 0u);
        std::vector<brp_engine::TopRow> rows(5 * size_t(top->rows));
        for (size_t r = 0; r < rows.size(); ++r) {
            const brp_bank_row& b = top->table[r];
            rows[r] = brp_engine::TopRow{float(b.power), float(b.period), float(b.tau), float(b.psi), b.bin, b.harmonics};
        }
        struct Buffers {
            std::vector<void*> held;
            void* get(size_t bytes) {
                void* p = nullptr;
                HIPC(hipMalloc(&p, std::max<size_t>(bytes, 4)));
                held.push_back(p);
                return p;
            }
            ~Buffers() { for (void* p : held) hipFree(p); }
        } d;
        auto* dslots = static_cast<uint32_t*>(d.get(sizeof(uint32_t) * slots.size()));
        auto* dparams = static_cast<brp_params*>(d.get(sizeof(brp_params) * params.size()));
        auto* dtable = static_cast<uint32_t*>(d.get(sizeof(uint32_t) * table.size()));
        auto* dorbit = static_cast<float4*>(d.get(sizeof(float4) * orbit.size()));
        auto* drows = static_cast<brp_engine::TopRow*>(d.get(sizeof(brp_engine::TopRow) * rows.size()));
        auto* derror = static_cast<uint32_t*>(d.get(sizeof(uint32_t)));
        auto* dlowest = static_cast<float*>(d.get(sizeof(float) * 5));
        HIPC(hipMemcpy(dslots, slots.data(), sizeof(uint32_t) * slots.size(), hipMemcpyHostToDevice));
        HIPC(hipMemcpy(dparams, params.data(), sizeof(brp_params) * params.size(), hipMemcpyHostToDevice));
        HIPC(hipMemcpy(dtable, table.data(), sizeof(uint32_t) * table.size(), hipMemcpyHostToDevice));
        HIPC(hipMemcpy(dorbit, orbit.data(), sizeof(float4) * orbit.size(), hipMemcpyHostToDevice));
        HIPC(hipMemcpy(drows, rows.data(), sizeof(brp_engine::TopRow) * rows.size(), hipMemcpyHostToDevice));
        HIPC(hipMemset(derror, 0, sizeof(uint32_t)));
        brp_engine::toplist_retain(nullptr, dslots, words, dparams, uint32_t(count), dtable, dorbit, 1, drows,
                                   top->rows, dlowest, derror);
        HIPC(hipGetLastError());
        uint32_t error = 0;
        HIPC(hipMemcpy(rows.data(), drows, sizeof(brp_engine::TopRow) * rows.size(), hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(&error, derror, sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (error & 2u) throw Failure(BRP_E_INVALID, "invalid compact candidate");
        if (error & 4u) throw Failure(BRP_E_RUNTIME, "device toplist: more than 4096 candidates of one level in one template");
        for (size_t r = 0; r < rows.size(); ++r) {
            const brp_engine::TopRow& v = rows[r];
            top->table[r] = brp_bank_row{double(v.power), double(v.period), double(v.tau), double(v.psi), 0.0,
                                         v.harmonics, v.bin};
        }
    });
}

int brp_hip_bank_search(brp_hip* h, const brp_orbit* orbits, size_t count, uint32_t samples, double sample_us,
                        double f0, double false_alarm, uint32_t window, brp_toplist* top, brp_params* templates_out,
                        uint32_t* cand_count, double* gpu_ms) {
    return guarded([&] {
        require(h && orbits && top, "null argument");
        // Parameters that do not depend on the toplist, once per template.  The plain engine
        // reads them on the host, so it takes brp_search_parameters whole.  On the fused
        // engine the host sets the per-template fields (src/workload.cpp:109-112, with the
        // host's sinf in S0) and the per-length ones once (window, F, H, thresholds,
        // src/workload.cpp:114-124), and the resampled lengths (resampling_steps) are left to
        // the device, in the launch stream of each batch; templates whose delay is too long
        // for the jump-started scan keep the host's scan.
        std::vector<brp_params> base(count);
        const bool device_steps = h->opt.engine == BRP_ENGINE_FUSED;
        brp_params limits{};
        for (size_t i = 0; i < count; ++i) {
            brp_params& p = base[i];
            if (!device_steps) {
                if (int rc = brp_search_parameters(h->base.n, samples, sample_us, &orbits[i], f0, false_alarm, window, &p))
                    throw Failure(rc, brp_last_error());
                continue;
            }
            if (int rc = brp::search_fields(h->base.n, samples, sample_us, orbits[i], f0, false_alarm, window, &p))
                throw Failure(rc, brp_last_error());
            if (!brp::steps_jump_start_safe(p)) {
                brp_params plain = p;
                plain.flags = 0;
                plain.s0 = brp::double_sine_s0(p);
                if (int rc = brp_resampling_steps(&plain)) throw Failure(rc, brp_last_error());
                if (int rc = brp_resampling_steps(&p)) throw Failure(rc, brp_last_error());
            }
            if (i == 0) {
                limits = p;
                if (int rc = brp::search_limits(sample_us, f0, false_alarm, window, &limits))
                    throw Failure(rc, brp_last_error());
            }
            p.window = limits.window;
            p.fundamental = limits.fundamental;
            p.harmonic = limits.harmonic;
            std::copy(limits.threshold, limits.threshold + 5, p.threshold);
        }
        brp_toplist* tops[1] = {top};
        bank_loop(*h, base, orbits, tops, 1, cand_count, templates_out, gpu_ms, device_steps);
    });
}

int brp_hip_bank_search_templates(brp_hip* h, const brp_params* templates, const brp_orbit* orbits, size_t count,
                                  brp_toplist* const* tops, size_t ntops, uint32_t* cand_count, double* gpu_ms) {
    return guarded([&] {
        require(h && templates && orbits && tops, "null argument");
        std::vector<brp_params> base(templates, templates + count);
        bank_loop(*h, base, orbits, tops, ntops, cand_count, nullptr, gpu_ms);
    });
}

}  // extern "C"
