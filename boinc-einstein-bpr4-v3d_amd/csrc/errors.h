// errors.h -- thread-local error message behind brp_last_error(), and the host science the
// engine shares with the C ABI.
#pragma once
#include <stdint.h>

#include <string>

#include "brp_hip.h"

namespace brp {
int fail(int code, const std::string& message);

// Host science shared by the C ABI and the engine (science.cpp).
int template_fields(uint32_t n, uint32_t unpadded, float tau, float period, float psi, float dt, brp_params* out);
float double_sine_s0(const brp_params& p);
int search_fields(uint32_t n, uint32_t samples, double sample_us, const brp_orbit& orbit, double f0,
                  double false_alarm, uint32_t window, brp_params* out);
int search_limits(double sample_us, double f0, double false_alarm, uint32_t window, brp_params* p);
// Whether resampling_steps may start its scan 64 samples above the FP64 solution: the
// FP32-vs-FP64 offset bound (tau/dt) x (2.5e-5 + 1e-6 (1 + |phase|)) is at most 15 samples
// and tau omega < 1/2.
bool steps_jump_start_safe(const brp_params& p);
}
