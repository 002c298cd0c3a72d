// hip.hpp -- the reference's C++ backend interface (include/brp/vulkan.hpp:6-29 and the
// science/workload declarations it uses) over the C ABI of brp_hip.h.  Header-only, so
// callers written against the reference's brp::GPU compile unchanged after swapping
// the include; errors surface as std::runtime_error with the library's message, as the
// reference's backend throws.
#pragma once
#include <array>
#include <complex>
#include <cstdint>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

#include "brp_hip.h"

namespace brp {

using Params = brp_params;        // include/brp/science.hpp:8-17
using Candidate = brp_candidate;  // include/brp/science.hpp:18
using Orbit = brp_orbit;          // include/brp/workload.hpp:13
using BankCandidate = brp_bank_row;
static_assert(sizeof(Params) == 108);
static_assert(sizeof(Candidate) == 16);

inline void check(int rc) {
    if (rc != BRP_OK) throw std::runtime_error(brp_last_error());
}

// src/science.cpp:46-65
inline Params parameters(uint32_t n, uint32_t unpadded, float tau, float period, float psi, float dt) {
    Params p{};
    check(brp_parameters(n, unpadded, tau, period, psi, dt, &p));
    return p;
}
inline void resampling_steps(Params& p) { check(brp_resampling_steps(&p)); }
inline float lut_sin(float phase) { return brp_lut_sin(phase); }
inline float sample_offset(uint32_t i, const Params& p) { return brp_sample_offset(i, &p); }

// src/workload.cpp:108-125 with the observation's (n, samples, sample_us).
inline Params search_parameters(uint32_t n, uint32_t samples, double sample_us, const Orbit& orbit,
                                double f0 = 400, double false_alarm = 0.08, uint32_t window = 1000) {
    Params p{};
    check(brp_search_parameters(n, samples, sample_us, &orbit, f0, false_alarm, window, &p));
    return p;
}

// include/brp/vulkan.hpp:6
struct Timings {
    double wall_ms, cpu_ms, submit_ms;
    std::array<double, 4> stage_ms;
    std::vector<std::array<double, 4>> template_stage_ms;
};

// include/brp/vulkan.hpp:7-11
struct Snapshot {
    std::vector<float> resampled, power, summed;
    std::vector<std::complex<float>> fft;
    std::vector<Candidate> candidates;
};

// include/brp/workload.hpp:16-23
class Toplist {
public:
    explicit Toplist(uint32_t rows = 100) {
        brp_toplist* t = nullptr;
        check(brp_toplist_create(rows, &t));
        impl_.reset(t);
    }
    void thresholds(Params& p) const { brp_toplist_thresholds(impl_.get(), &p); }
    void retain(std::span<const Candidate> c, const Orbit& orbit, const Params& p) {
        check(brp_toplist_retain(impl_.get(), c.data(), c.size(), &orbit, &p));
    }
    std::vector<BankCandidate> final() const {
        std::vector<BankCandidate> out(100000);
        out.resize(brp_toplist_final(impl_.get(), out.data(), out.size()));
        return out;
    }
    brp_toplist* handle() { return impl_.get(); }

private:
    struct Free { void operator()(brp_toplist* t) const { brp_toplist_destroy(t); } };
    std::unique_ptr<brp_toplist, Free> impl_;
};

// include/brp/vulkan.hpp:12-29
class GPU {
public:
    GPU(const Params& p, std::span<const float> input, uint32_t workgroup, uint32_t tile, uint32_t batch,
        bool software = false, bool stockham = false, bool radix4 = false, bool packed = false, bool scatter = false,
        bool radix8 = false, bool fused_power = false, bool radix8_store = false,
        uint32_t engine = BRP_ENGINE_PLAIN, bool graph = false, bool snapshot = false)
        : p_(p), batch_(batch) {
        brp_options o{};
        o.workgroup = workgroup; o.tile = tile; o.batch = batch; o.software = software; o.stockham = stockham;
        o.radix4 = radix4; o.packed = packed; o.scatter = scatter; o.radix8 = radix8; o.fused_power = fused_power;
        o.radix8_store = radix8_store; o.engine = engine; o.graph = graph; o.snapshot = snapshot;
        brp_hip* h = nullptr;
        check(brp_hip_create(&p, input.data(), input.size(), &o, &h));
        impl_.reset(h);
        templates_.assign(1, p);
    }
    // Every backend option (brp_options: the reference's switches plus engine, group,
    // split, lanes and the other extensions).
    GPU(const Params& p, std::span<const float> input, const brp_options& o) : p_(p), batch_(o.batch) {
        brp_hip* h = nullptr;
        check(brp_hip_create(&p, input.data(), input.size(), &o, &h));
        impl_.reset(h);
        templates_.assign(1, p);
    }
    GPU(const GPU&) = delete;
    GPU& operator=(const GPU&) = delete;

    Timings run() {
        brp_timings t{};
        std::vector<double> per(4 * templates_.size());
        check(brp_hip_run(impl_.get(), &t, per.data()));
        Timings out{t.wall_ms, t.cpu_ms, t.submit_ms, {t.stage_ms[0], t.stage_ms[1], t.stage_ms[2], t.stage_ms[3]}, {}};
        out.template_stage_ms.resize(templates_.size());
        for (size_t r = 0; r < templates_.size(); ++r)
            for (int s = 0; s < 4; ++s) out.template_stage_ms[r][s] = per[4 * r + s];
        return out;
    }
    void set_template(const Params& p) { set_templates(std::span<const Params>(&p, 1)); }
    void set_templates(std::span<const Params> ps) {
        check(brp_hip_set_templates(impl_.get(), ps.data(), ps.size()));
        templates_.assign(ps.begin(), ps.end());
    }
    Snapshot snapshot() const {
        const uint32_t slot = uint32_t(templates_.size() - 1);
        const Params& p = templates_.back();
        Snapshot s;
        s.resampled.resize(p.n);
        s.power.resize(p.n / 2);
        s.summed.resize(5ull * p.fundamental);
        s.fft.resize(p.n / 2 + 1);
        check(brp_hip_snapshot(impl_.get(), slot, s.resampled.data(), nullptr,
                               reinterpret_cast<float*>(s.fft.data()), s.power.data(), s.summed.data()));
        s.candidates = candidates(slot);
        return s;
    }
    std::vector<Candidate> candidates(uint32_t slot = UINT32_MAX) const {
        if (slot == UINT32_MAX) slot = uint32_t(templates_.size() - 1);
        size_t n = 0;
        check(brp_hip_candidates(impl_.get(), slot, nullptr, 0, &n));
        std::vector<Candidate> out(n);
        check(brp_hip_candidates(impl_.get(), slot, out.data(), out.size(), &n));
        return out;
    }
    std::string device_info() const {
        std::string s(brp_hip_device_info(impl_.get(), nullptr, 0) + 1, '\0');
        brp_hip_device_info(impl_.get(), s.data(), s.size());
        s.pop_back();
        return s;
    }
    uint64_t allocated_bytes() const { return brp_hip_allocated_bytes(impl_.get()); }
    std::string profile_counters() {
        std::string s(brp_hip_profile_counters(impl_.get(), nullptr, 0) + 1, '\0');
        brp_hip_profile_counters(impl_.get(), s.data(), s.size());
        s.pop_back();
        return s;
    }
    brp_hip* handle() { return impl_.get(); }

    // Batched extension (fused engine): another observation on the same handle; templates
    // name it by (n, unpadded).
    void add_series(const Params& p, std::span<const float> input) {
        check(brp_hip_add_series(impl_.get(), &p, input.data(), input.size()));
    }
    // src/workunit_main.cpp:77-136 over prepared templates of any registered series, one
    // toplist per series (registration order); returns the candidate count per template.
    std::vector<uint32_t> bank_search(std::span<const Params> templates, std::span<const brp_orbit> orbits,
                                      std::span<brp_toplist* const> tops) {
        if (templates.size() != orbits.size()) throw std::runtime_error("templates and orbits differ in length");
        std::vector<uint32_t> counts(templates.size());
        double gpu_ms = 0;
        check(brp_hip_bank_search_templates(impl_.get(), templates.data(), orbits.data(), templates.size(),
                                            tops.data(), tops.size(), counts.data(), &gpu_ms));
        return counts;
    }

private:
    struct Free { void operator()(brp_hip* h) const { brp_hip_destroy(h); } };
    std::unique_ptr<brp_hip, Free> impl_;
    Params p_;
    uint32_t batch_;
    std::vector<Params> templates_;
};

// load_observation's decode + preprocess (src/workload.cpp:29-95) on the device, for the
// sample block of a DD file; returns the preprocessed series (samples floats).
inline std::vector<float> preprocess(std::span<const uint8_t> block, size_t samples, uint32_t bits, double scale,
                                     uint32_t n, uint32_t window = 1000, float rfi_ratio = 30.0f,
                                     uint32_t* rfi_bins = nullptr) {
    std::vector<float> out(samples);
    check(brp_hip_preprocess(block.data(), samples, bits, scale, n, window, rfi_ratio, out.data(), nullptr, rfi_bins));
    return out;
}

// The same with the reference's zap list (src/workload.cpp:76-86): (lower, upper) Hz
// pairs flattened; decoded (may be null) receives the decoded samples.
inline std::vector<float> preprocess(std::span<const uint8_t> block, size_t samples, uint32_t bits, double scale,
                                     uint32_t n, uint32_t window, float rfi_ratio, std::span<const double> zap,
                                     double sample_us, double padding, std::vector<float>* decoded = nullptr,
                                     uint32_t* rfi_bins = nullptr) {
    std::vector<float> out(samples);
    if (decoded) decoded->resize(samples);
    check(brp_hip_preprocess_zap(block.data(), samples, bits, scale, n, window, rfi_ratio, zap.data(), zap.size() / 2,
                                 sample_us, padding, out.data(), decoded ? decoded->data() : nullptr, rfi_bins));
    return out;
}

}  // namespace brp
