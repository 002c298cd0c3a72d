// hip.hpp -- drop-in for the reference's include/brp/vulkan.hpp: the same brp::Timings,
// brp::Snapshot and brp::GPU (include/brp/vulkan.hpp:6-29) over the C ABI of brp_hip.h.
//
// A caller of the reference swaps `#include "brp/vulkan.hpp"` for `#include "brp/hip.hpp"`,
// adds this include directory and links libbrp_hip.so in place of brp_vulkan.  Like
// vulkan.hpp, this header includes "brp/science.hpp": inside the reference's tree that is
// the reference's own science library (Params, Candidate, the CPU oracle), which the caller
// keeps linking, and its workload library keeps providing Orbit, Toplist and
// search_parameters.  Without the reference's headers it is this library's mirror,
// brp/hip_science.hpp.  Params and Candidate cross the ABI as brp_params and brp_candidate,
// whose layouts are checked member by member below.  Errors surface as std::runtime_error
// with the library's message, as the reference's backend throws.
#pragma once
#if __has_include("brp/science.hpp")
#include "brp/science.hpp"
#else
#include "brp/hip_science.hpp"
#endif

#include <algorithm>
#include <array>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <span>
#include <string>
#include <type_traits>
#include <vector>

#include "brp/hip_error.hpp"
#include "brp_hip.h"

// brp_vulkan exports this definition to its users (CMakeLists.txt:39); its twiddle tables
// here are always dense.
#ifndef BRP_STRIDED_TWIDDLES
#define BRP_STRIDED_TWIDDLES 0
#endif

namespace brp {

namespace hip_detail {
// brp::Params and brp_params: the same 108-byte block (include/brp/science.hpp:8-17).
static_assert(std::is_standard_layout_v<Params> && sizeof(Params) == sizeof(brp_params));
#define BRP_HIP_SAME(T, C, f) static_assert(offsetof(T, f) == offsetof(C, f), #f)
BRP_HIP_SAME(Params, brp_params, n); BRP_HIP_SAME(Params, brp_params, unpadded);
BRP_HIP_SAME(Params, brp_params, steps); BRP_HIP_SAME(Params, brp_params, logn);
BRP_HIP_SAME(Params, brp_params, tau); BRP_HIP_SAME(Params, brp_params, omega);
BRP_HIP_SAME(Params, brp_params, psi); BRP_HIP_SAME(Params, brp_params, dt);
BRP_HIP_SAME(Params, brp_params, step_inv); BRP_HIP_SAME(Params, brp_params, s0);
BRP_HIP_SAME(Params, brp_params, fundamental); BRP_HIP_SAME(Params, brp_params, harmonic);
BRP_HIP_SAME(Params, brp_params, window); BRP_HIP_SAME(Params, brp_params, stage);
BRP_HIP_SAME(Params, brp_params, threshold); BRP_HIP_SAME(Params, brp_params, capacity);
BRP_HIP_SAME(Params, brp_params, fft_n); BRP_HIP_SAME(Params, brp_params, fft_base);
BRP_HIP_SAME(Params, brp_params, fft_radix); BRP_HIP_SAME(Params, brp_params, flags);
BRP_HIP_SAME(Params, brp_params, twiddle3); BRP_HIP_SAME(Params, brp_params, twiddle_real);
BRP_HIP_SAME(Params, brp_params, result_slot);
// brp::Candidate and brp_candidate (include/brp/science.hpp:18).
static_assert(std::is_standard_layout_v<Candidate> && sizeof(Candidate) == sizeof(brp_candidate));
BRP_HIP_SAME(Candidate, brp_candidate, bin); BRP_HIP_SAME(Candidate, brp_candidate, level);
BRP_HIP_SAME(Candidate, brp_candidate, power); BRP_HIP_SAME(Candidate, brp_candidate, reserved);
#undef BRP_HIP_SAME

inline const brp_params* c_params(const Params* p) { return reinterpret_cast<const brp_params*>(p); }
inline brp_candidate* c_candidates(Candidate* c) { return reinterpret_cast<brp_candidate*>(c); }
}  // namespace hip_detail

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

// include/brp/vulkan.hpp:12-29.  The reference's constructor runs the plain engine (the
// reference's launch sequence, one kernel per stage); the brp_options constructor reaches
// every backend option, the batched fused engine among them.
class GPU {
public:
    GPU(const Params& p, std::span<const float> input, uint32_t workgroup, uint32_t tile, uint32_t batch,
        bool software = false, bool stockham = false, bool radix4 = false, bool packed = false, bool scatter = false,
        bool radix8 = false, bool fused_power = false, bool radix8_store = false)
        : GPU(p, input, reference_options(workgroup, tile, batch, software, stockham, radix4, packed, scatter, radix8,
                                          fused_power, radix8_store)) {}
    GPU(const Params& p, std::span<const float> input, const brp_options& options) {
        brp_hip* h = nullptr;
        hip_detail::check(brp_hip_create(hip_detail::c_params(&p), input.data(), input.size(), &options, &h));
        handle_.reset(h);
        current_.assign(1, p);
    }
    ~GPU() = default;
    GPU(const GPU&) = delete;
    GPU& operator=(const GPU&) = delete;

    Timings run() {
        brp_timings t{};
        std::vector<double> per(4 * current_.size());
        hip_detail::check(brp_hip_run(handle_.get(), &t, per.data()));
        Timings out{t.wall_ms, t.cpu_ms, t.submit_ms, {t.stage_ms[0], t.stage_ms[1], t.stage_ms[2], t.stage_ms[3]}, {}};
        out.template_stage_ms.resize(current_.size());
        for (size_t slot = 0; slot < current_.size(); ++slot)
            for (size_t s = 0; s < 4; ++s) out.template_stage_ms[slot][s] = per[4 * slot + s];
        return out;
    }
    void set_template(const Params& p) { set_templates(std::span<const Params>(&p, 1)); }
    void set_templates(std::span<const Params> ps) {
        hip_detail::check(brp_hip_set_templates(handle_.get(), hip_detail::c_params(ps.data()), ps.size()));
        current_.assign(ps.begin(), ps.end());
    }
    // The last template of the batch: its resampled series with the mean pad, spectrum,
    // power, harmonic sums and candidates (src/vulkan.cpp:319-328).
    Snapshot snapshot() const {
        const uint32_t slot = uint32_t(current_.size() - 1);
        const Params& p = current_.back();
        Snapshot s;
        s.resampled.resize(p.n);
        s.fft.resize(p.n / 2 + 1);
        s.power.resize(p.n / 2);
        s.summed.resize(5 * size_t(p.fundamental));
        hip_detail::check(brp_hip_snapshot(handle_.get(), slot, s.resampled.data(), nullptr,
                                           reinterpret_cast<float*>(s.fft.data()), s.power.data(), s.summed.data()));
        s.candidates = candidates(slot);
        return s;
    }
    std::vector<Candidate> candidates(uint32_t slot = UINT32_MAX) const {
        if (slot == UINT32_MAX) slot = uint32_t(current_.size() - 1);
        size_t count = 0;
        hip_detail::check(brp_hip_candidates(handle_.get(), slot, nullptr, 0, &count));
        std::vector<Candidate> out(count);
        hip_detail::check(brp_hip_candidates(handle_.get(), slot, hip_detail::c_candidates(out.data()), out.size(),
                                             &count));
        return out;
    }
    std::string device_info() const { return text(brp_hip_device_info); }
    uint64_t allocated_bytes() const { return brp_hip_allocated_bytes(handle_.get()); }
    // Collected on every call (one replay of the batch per counter set), so the text is read in
    // one call into a buffer that is grown and read again only when it was too small.
    std::string profile_counters() {
        std::string s(size_t(1) << 20, '\0');
        size_t n = brp_hip_profile_counters(handle_.get(), s.data(), s.size());
        if (n >= s.size()) {
            s.assign(n + 1, '\0');
            n = brp_hip_profile_counters(handle_.get(), s.data(), s.size());
        }
        s.resize(std::min(n, s.size() - 1));
        return s;
    }

    // Extensions (fused engine): the C handle, another observation on it (templates name
    // their series by (n, unpadded)), and the bank loop over prepared templates of any
    // registered series with one C toplist per series; returns each template's candidate
    // count.
    brp_hip* handle() const { return handle_.get(); }
    void add_series(const Params& p, std::span<const float> input) {
        hip_detail::check(brp_hip_add_series(handle_.get(), hip_detail::c_params(&p), input.data(), input.size()));
    }
    std::vector<uint32_t> bank_search(std::span<const Params> templates, std::span<const brp_orbit> orbits,
                                      std::span<brp_toplist* const> tops) {
        if (templates.size() != orbits.size()) throw std::runtime_error("templates and orbits differ in length");
        std::vector<uint32_t> counts(templates.size());
        double gpu_ms = 0;
        hip_detail::check(brp_hip_bank_search_templates(handle_.get(), hip_detail::c_params(templates.data()),
                                                        orbits.data(), templates.size(), tops.data(), tops.size(),
                                                        counts.data(), &gpu_ms));
        return counts;
    }

private:
    static brp_options reference_options(uint32_t workgroup, uint32_t tile, uint32_t batch, bool software,
                                         bool stockham, bool radix4, bool packed, bool scatter, bool radix8,
                                         bool fused_power, bool radix8_store) {
        brp_options o{};
        o.workgroup = workgroup;
        o.tile = tile;
        o.batch = batch;
        o.software = software;
        o.stockham = stockham;
        o.radix4 = radix4;
        o.packed = packed;
        o.scatter = scatter;
        o.radix8 = radix8;
        o.fused_power = fused_power;
        o.radix8_store = radix8_store;
        o.engine = BRP_ENGINE_PLAIN;
        return o;
    }
    template <class F>
    std::string text(F&& read) const {
        std::string s(read(handle_.get(), nullptr, 0) + 1, '\0');
        read(handle_.get(), s.data(), s.size());
        s.pop_back();
        return s;
    }
    struct Destroy { void operator()(brp_hip* h) const { brp_hip_destroy(h); } };
    std::unique_ptr<brp_hip, Destroy> handle_;
    std::vector<Params> current_;
};

}  // namespace brp
