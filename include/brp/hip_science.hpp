// hip_science.hpp -- this library's host science and bank declarations in the reference's
// C++ shapes, over the C ABI of brp_hip.h.
//
// It stands in for the reference's include/brp/science.hpp:8-38 and
// include/brp/workload.hpp:6-24 where those are absent, that is, in builds against this
// library alone (its own drivers under cli/).  brp/hip.hpp includes the reference's
// brp/science.hpp instead when the include path has one, so inside the reference's tree the
// reference's science and workload libraries keep providing these names and nothing here is
// declared twice.  Only what this library implements is declared: the parameter set-up
// (src/science.cpp:20-65), the bank semantics (src/workload.cpp:108-165) and the rel-L2
// metric (src/science.cpp:121-127); the reference's FFTW oracle, historical adapters and
// CPU benchmark are test infrastructure here (oracle/).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <filesystem>
#include <span>
#include <stdexcept>
#include <vector>

#include "brp/hip_error.hpp"
#include "brp_hip.h"

namespace brp {

// include/brp/science.hpp:8-17 (the 108-byte ABI block; the same member sequence as
// brp_params, checked member by member in brp/hip.hpp).
struct Params {
    uint32_t n, unpadded, steps, logn;
    float tau, omega, psi, dt, step_inv, s0;
    uint32_t fundamental, harmonic, window, stage;
    float threshold[5];
    uint32_t capacity;
    uint32_t fft_n, fft_base, fft_radix, flags, twiddle3, twiddle_real;
    uint32_t result_slot;
};
static_assert(sizeof(Params) == 108);

// include/brp/science.hpp:18
struct Candidate { uint32_t bin, level; float power; uint32_t reserved; };

namespace hip_detail {
inline const brp_params* as_c(const Params* p) { return reinterpret_cast<const brp_params*>(p); }
inline brp_params* as_c(Params* p) { return reinterpret_cast<brp_params*>(p); }
}  // namespace hip_detail

inline float lut_sin(float phase) { return brp_lut_sin(phase); }
inline float sample_offset(uint32_t i, const Params& p) { return brp_sample_offset(i, hip_detail::as_c(&p)); }
inline void resampling_steps(Params& p) { hip_detail::check(brp_resampling_steps(hip_detail::as_c(&p))); }
inline Params parameters(uint32_t n, uint32_t unpadded, float tau, float period, float psi, float dt) {
    Params p{};
    hip_detail::check(brp_parameters(n, unpadded, tau, period, psi, dt, hip_detail::as_c(&p)));
    return p;
}

// |actual - expected| / |expected| in double; infinite when a value is not finite.
inline double relative_l2(std::span<const float> actual, std::span<const float> expected) {
    if (actual.size() != expected.size()) throw std::runtime_error("comparison length mismatch");
    double diff = 0, norm = 0;
    for (size_t i = 0; i < actual.size(); ++i) {
        if (!std::isfinite(actual[i]) || !std::isfinite(expected[i])) return INFINITY;
        const double d = double(actual[i]) - double(expected[i]);
        diff += d * d;
        norm += double(expected[i]) * double(expected[i]);
    }
    return std::sqrt(diff / std::max(norm, 1e-30));
}

// include/brp/workload.hpp:6-10; bits and block keep the undecoded samples for the device
// decode (this library's preprocessing runs on the GPU).
struct Observation {
    double sample_us{}, observation_s{}, dm{}, scale{}, padding{};
    uint32_t samples{}, n{}, bits{};
    std::vector<uint8_t> block;
    std::vector<float> decoded, preprocessed;
};

// include/brp/workload.hpp:13
struct Orbit { double period, tau, psi; };

// src/workload.cpp:108-125
inline Params search_parameters(const Observation& o, const Orbit& orbit, double f0 = 400, double false_alarm = 0.08,
                                uint32_t window = 1000) {
    Params p{};
    const brp_orbit c{orbit.period, orbit.tau, orbit.psi};
    hip_detail::check(brp_search_parameters(o.n, o.samples, o.sample_us, &c, f0, false_alarm, window,
                                            hip_detail::as_c(&p)));
    return p;
}

// include/brp/workload.hpp:15
struct BankCandidate {
    double power{}, period{}, tau{}, psi{}, significance{};
    uint32_t harmonics{}, bin{};
};

// include/brp/workload.hpp:16-23, with the rows per level as a parameter (reference: 100).
class Toplist {
public:
    explicit Toplist(uint32_t rows = 100) : rows_(rows) {
        brp_toplist* t = nullptr;
        hip_detail::check(brp_toplist_create(rows, &t));
        table_.reset(t);
    }
    void thresholds(Params& p) const { brp_toplist_thresholds(table_.get(), hip_detail::as_c(&p)); }
    void retain(std::span<const Candidate> found, const Orbit& orbit, const Params& p) {
        const brp_orbit c{orbit.period, orbit.tau, orbit.psi};
        hip_detail::check(brp_toplist_retain(table_.get(), reinterpret_cast<const brp_candidate*>(found.data()),
                                             found.size(), &c, hip_detail::as_c(&p)));
    }
    std::vector<BankCandidate> final() const {
        std::vector<BankCandidate> rows(rows_);
        rows.resize(brp_toplist_final(table_.get(), reinterpret_cast<brp_bank_row*>(rows.data()), rows.size()));
        return rows;
    }
    brp_toplist* handle() const { return table_.get(); }

private:
    struct Destroy { void operator()(brp_toplist* t) const { brp_toplist_destroy(t); } };
    uint32_t rows_;
    std::unique_ptr<brp_toplist, Destroy> table_;
};

// src/workload.cpp:160-165
inline void write_candidates(const std::filesystem::path& path, std::span<const BankCandidate> rows,
                             double observation_s) {
    hip_detail::check(brp_write_candidates(path.c_str(), reinterpret_cast<const brp_bank_row*>(rows.data()),
                                           rows.size(), observation_s));
}

static_assert(sizeof(Candidate) == sizeof(brp_candidate) && sizeof(Orbit) == sizeof(brp_orbit) &&
              sizeof(BankCandidate) == sizeof(brp_bank_row));
static_assert(offsetof(Candidate, power) == offsetof(brp_candidate, power) &&
              offsetof(BankCandidate, significance) == offsetof(brp_bank_row, significance) &&
              offsetof(BankCandidate, harmonics) == offsetof(brp_bank_row, harmonics) &&
              offsetof(BankCandidate, bin) == offsetof(brp_bank_row, bin));

}  // namespace brp
