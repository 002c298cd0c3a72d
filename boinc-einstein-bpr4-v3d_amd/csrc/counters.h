// counters.h -- in-process hardware counters (the role of the reference's Vulkan performance
// queries in GPU::profile_counters, src/vulkan.cpp:346-406) over rocprofiler-sdk's dispatch
// counting service.
//
// Collection has to be armed before the ROCm runtime starts in the process: with
// BRP_HIP_COUNTERS=1 in the environment, the library's load-time initialiser loads
// rocprofiler-sdk and registers itself as a tool.  Without it (or when another tool such as
// rocprofv3 already holds the configuration) collection is unavailable and says why.
#pragma once
#include <stdint.h>

#include <functional>
#include <map>
#include <string>
#include <vector>

namespace brp_counters {

// "" when collection can run, else the reason it cannot.
std::string unavailable_reason();

struct Dispatch {
    uint64_t id = 0;
    std::string kernel;                    // kernel symbol name
    uint32_t grid[3] = {0, 0, 0};          // work-items
    uint32_t workgroup[3] = {0, 0, 0};
    uint64_t start_ns = 0, end_ns = 0;     // as timed by the counting service
    std::map<std::string, double> values;  // counter name -> sum over its instances
};

// Runs `body` (which enqueues work and waits for it) with `counters` collected on every kernel
// dispatch the process issues meanwhile (other threads' included); returns the dispatches in
// issue order.  Counters the device does not support are left out (`missing` lists them).
// Calls from several threads run one after another.  Throws std::runtime_error when
// unavailable or when the records do not arrive.
std::vector<Dispatch> collect(const std::vector<std::string>& counters, const std::function<void()>& body,
                              std::vector<std::string>* missing = nullptr);

}  // namespace brp_counters
