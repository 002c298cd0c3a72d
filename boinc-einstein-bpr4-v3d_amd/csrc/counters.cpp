// counters.cpp -- in-process hardware counters over rocprofiler-sdk (see counters.h).
//
// rocprofiler-sdk is loaded with dlopen only when BRP_HIP_COUNTERS=1, so the library neither
// links it nor registers a tool otherwise.  The tool holds two contexts: one traces kernel
// symbol registration (kernel id -> name) from the start, the other runs the dispatch counting
// service and is started only inside collect(); its dispatch callback hands the requested
// counter configuration to dispatches issued while collecting and its record callback sums each
// counter's instances per dispatch.
#include "counters.h"

#include <dlfcn.h>
#include <rocprofiler-sdk/callback_tracing.h>
#include <rocprofiler-sdk/context.h>
#include <rocprofiler-sdk/counter_config.h>
#include <rocprofiler-sdk/counters.h>
#include <rocprofiler-sdk/dispatch_counting_service.h>
#include <rocprofiler-sdk/fwd.h>
#include <rocprofiler-sdk/registration.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <unordered_map>

namespace brp_counters {
namespace {

struct Api {
    decltype(&rocprofiler_force_configure) force_configure = nullptr;
    decltype(&rocprofiler_create_context) create_context = nullptr;
    decltype(&rocprofiler_start_context) start_context = nullptr;
    decltype(&rocprofiler_stop_context) stop_context = nullptr;
    decltype(&rocprofiler_configure_callback_dispatch_counting_service) configure_counting = nullptr;
    decltype(&rocprofiler_configure_callback_tracing_service) configure_tracing = nullptr;
    decltype(&rocprofiler_iterate_agent_supported_counters) iterate_counters = nullptr;
    decltype(&rocprofiler_query_counter_info) counter_info = nullptr;
    decltype(&rocprofiler_query_record_counter_id) record_counter_id = nullptr;
    decltype(&rocprofiler_create_counter_config) create_config = nullptr;
};

struct State {
    std::mutex mu;
    std::mutex collect_mu;  // one collection at a time
    Api api;
    std::string reason = "in-process counters need BRP_HIP_COUNTERS=1 in the environment when the library loads";
    bool initialized = false;  // the tool's contexts exist
    rocprofiler_context_id_t counting{}, tracing{};
    std::unordered_map<uint64_t, std::string> kernel_names;
    std::unordered_map<uint64_t, std::string> counter_names;                // counter id -> name
    std::map<uint64_t, std::map<std::string, uint64_t>> agent_counters;     // agent -> name -> id
    std::map<std::pair<uint64_t, std::string>, uint64_t> configs;           // (agent, set) -> config
    std::map<std::pair<uint64_t, std::string>, std::vector<std::string>> config_missing;  // its absent names
    std::atomic<bool> collecting{false};
    std::vector<std::string> wanted;
    std::string wanted_key;
    std::vector<std::string> missing;
    std::map<uint64_t, Dispatch> dispatches;
    uint64_t issued = 0, completed = 0;
};

State& state() {
    static State s;
    return s;
}

// The agent's counter ids by name (under the state lock).
const std::map<std::string, uint64_t>& agent_counters(State& s, rocprofiler_agent_id_t agent) {
    auto it = s.agent_counters.find(agent.handle);
    if (it != s.agent_counters.end()) return it->second;
    auto& table = s.agent_counters[agent.handle];
    s.api.iterate_counters(
        agent,
        [](rocprofiler_agent_id_t, rocprofiler_counter_id_t* ids, size_t n, void* user) -> rocprofiler_status_t {
            auto& st = state();
            auto& tb = *static_cast<std::map<std::string, uint64_t>*>(user);
            for (size_t i = 0; i < n; ++i) {
                rocprofiler_counter_info_v0_t info{};
                if (st.api.counter_info(ids[i], ROCPROFILER_COUNTER_INFO_VERSION_0, &info) == ROCPROFILER_STATUS_SUCCESS &&
                    info.name) {
                    tb[info.name] = ids[i].handle;
                    st.counter_names[ids[i].handle] = info.name;
                }
            }
            return ROCPROFILER_STATUS_SUCCESS;
        },
        &table);
    return table;
}

void on_dispatch(rocprofiler_dispatch_counting_service_data_t data, rocprofiler_counter_config_id_t* config,
                 rocprofiler_user_data_t*, void*) {
    auto& s = state();
    if (!s.collecting.load()) return;
    std::lock_guard<std::mutex> lock(s.mu);
    const auto agent = data.dispatch_info.agent_id;
    const auto key = std::make_pair(agent.handle, s.wanted_key);
    auto it = s.configs.find(key);
    if (it == s.configs.end()) {
        const auto& table = agent_counters(s, agent);
        std::vector<rocprofiler_counter_id_t> ids;
        auto& absent = s.config_missing[key];
        for (const auto& name : s.wanted) {
            auto c = table.find(name);
            if (c == table.end()) {
                absent.push_back(name);
                continue;
            }
            rocprofiler_counter_id_t id{};
            id.handle = c->second;
            ids.push_back(id);
        }
        rocprofiler_counter_config_id_t cfg{};
        if (ids.empty() || s.api.create_config(agent, ids.data(), ids.size(), &cfg) != ROCPROFILER_STATUS_SUCCESS)
            cfg.handle = 0;
        it = s.configs.emplace(key, cfg.handle).first;
    }
    for (const auto& name : s.config_missing[key])
        if (std::find(s.missing.begin(), s.missing.end(), name) == s.missing.end()) s.missing.push_back(name);
    if (!it->second) return;
    config->handle = it->second;
    Dispatch d;
    d.id = data.dispatch_info.dispatch_id;
    auto name = s.kernel_names.find(data.dispatch_info.kernel_id);
    d.kernel = name == s.kernel_names.end() ? "kernel " + std::to_string(data.dispatch_info.kernel_id) : name->second;
    const auto& wg = data.dispatch_info.workgroup_size;
    const auto& gr = data.dispatch_info.grid_size;
    d.workgroup[0] = wg.x; d.workgroup[1] = wg.y; d.workgroup[2] = wg.z;
    d.grid[0] = gr.x; d.grid[1] = gr.y; d.grid[2] = gr.z;
    s.dispatches[d.id] = d;
    ++s.issued;
}

void on_record(rocprofiler_dispatch_counting_service_data_t data, rocprofiler_counter_record_t* records, size_t count,
               rocprofiler_user_data_t, void*) {
    auto& s = state();
    std::lock_guard<std::mutex> lock(s.mu);
    auto it = s.dispatches.find(data.dispatch_info.dispatch_id);
    if (it == s.dispatches.end()) return;
    Dispatch& d = it->second;
    d.start_ns = data.start_timestamp;
    d.end_ns = data.end_timestamp;
    for (size_t i = 0; i < count; ++i) {
        rocprofiler_counter_id_t id{};
        if (s.api.record_counter_id(records[i].id, &id) != ROCPROFILER_STATUS_SUCCESS) continue;
        auto n = s.counter_names.find(id.handle);
        if (n != s.counter_names.end()) d.values[n->second] += records[i].counter_value;
    }
    ++s.completed;
}

void on_code_object(rocprofiler_callback_tracing_record_t record, rocprofiler_user_data_t*, void*) {
    if (record.kind != ROCPROFILER_CALLBACK_TRACING_CODE_OBJECT ||
        record.operation != ROCPROFILER_CODE_OBJECT_DEVICE_KERNEL_SYMBOL_REGISTER ||
        record.phase != ROCPROFILER_CALLBACK_PHASE_LOAD)
        return;
    const auto* sym = static_cast<rocprofiler_callback_tracing_code_object_kernel_symbol_register_data_t*>(record.payload);
    auto& s = state();
    std::lock_guard<std::mutex> lock(s.mu);
    if (sym && sym->kernel_name) s.kernel_names[sym->kernel_id] = sym->kernel_name;
}

int tool_init(rocprofiler_client_finalize_t, void*) {
    auto& s = state();
    std::lock_guard<std::mutex> lock(s.mu);
    const rocprofiler_tracing_operation_t ops[] = {ROCPROFILER_CODE_OBJECT_DEVICE_KERNEL_SYMBOL_REGISTER};
    if (s.api.create_context(&s.tracing) != ROCPROFILER_STATUS_SUCCESS ||
        s.api.create_context(&s.counting) != ROCPROFILER_STATUS_SUCCESS ||
        s.api.configure_tracing(s.tracing, ROCPROFILER_CALLBACK_TRACING_CODE_OBJECT, ops, 1, on_code_object, nullptr) !=
            ROCPROFILER_STATUS_SUCCESS ||
        s.api.configure_counting(s.counting, on_dispatch, nullptr, on_record, nullptr) != ROCPROFILER_STATUS_SUCCESS ||
        s.api.start_context(s.tracing) != ROCPROFILER_STATUS_SUCCESS) {
        s.reason = "rocprofiler-sdk refused the counting or tracing configuration";
        return -1;
    }
    s.initialized = true;
    s.reason.clear();
    return 0;
}

void tool_fini(void*) {}

rocprofiler_tool_configure_result_t* configure(uint32_t, const char*, uint32_t, rocprofiler_client_id_t* id) {
    if (id) id->name = "libbrp_hip counters";
    static rocprofiler_tool_configure_result_t cfg{sizeof(rocprofiler_tool_configure_result_t), tool_init, tool_fini,
                                                   nullptr};
    return &cfg;
}

// Loads rocprofiler-sdk and registers the tool when BRP_HIP_COUNTERS=1 (at library load, before
// the runtime starts).
__attribute__((constructor)) void arm() {
    const char* want = std::getenv("BRP_HIP_COUNTERS");
    if (!want || std::strcmp(want, "1") != 0) return;
    auto& s = state();
    void* lib = dlopen("librocprofiler-sdk.so.1", RTLD_NOW | RTLD_GLOBAL);
    if (!lib) lib = dlopen("librocprofiler-sdk.so", RTLD_NOW | RTLD_GLOBAL);
    if (!lib) {
        s.reason = "librocprofiler-sdk.so could not be loaded";
        return;
    }
    Api& a = s.api;
    bool ok = true;
    auto get = [&](auto& fn, const char* name) {
        fn = reinterpret_cast<std::remove_reference_t<decltype(fn)>>(dlsym(lib, name));
        ok = ok && fn;
    };
    get(a.force_configure, "rocprofiler_force_configure");
    get(a.create_context, "rocprofiler_create_context");
    get(a.start_context, "rocprofiler_start_context");
    get(a.stop_context, "rocprofiler_stop_context");
    get(a.configure_counting, "rocprofiler_configure_callback_dispatch_counting_service");
    get(a.configure_tracing, "rocprofiler_configure_callback_tracing_service");
    get(a.iterate_counters, "rocprofiler_iterate_agent_supported_counters");
    get(a.counter_info, "rocprofiler_query_counter_info");
    get(a.record_counter_id, "rocprofiler_query_record_counter_id");
    get(a.create_config, "rocprofiler_create_counter_config");
    if (!ok) {
        s.reason = "librocprofiler-sdk.so lacks the dispatch counting interface";
        return;
    }
    if (a.force_configure(&configure) != ROCPROFILER_STATUS_SUCCESS)
        s.reason = "rocprofiler-sdk is already configured in this process (another tool such as rocprofv3, or the "
                   "runtime started before the library loaded)";
}

}  // namespace

std::string unavailable_reason() {
    auto& s = state();
    std::lock_guard<std::mutex> lock(s.mu);
    return s.initialized ? std::string() : s.reason;
}

std::vector<Dispatch> collect(const std::vector<std::string>& counters, const std::function<void()>& body,
                              std::vector<std::string>* missing) {
    auto& s = state();
    std::lock_guard<std::mutex> one(s.collect_mu);
    {
        std::lock_guard<std::mutex> lock(s.mu);
        if (!s.initialized) throw std::runtime_error(s.reason);
        s.wanted = counters;
        s.wanted_key.clear();
        for (const auto& c : counters) s.wanted_key += c + ",";
        s.dispatches.clear();
        s.missing.clear();
        s.issued = s.completed = 0;
        if (s.api.start_context(s.counting) != ROCPROFILER_STATUS_SUCCESS)
            throw std::runtime_error("rocprofiler-sdk could not start the counting context");
        s.collecting = true;
    }
    auto stop = [&] {
        s.collecting = false;
        s.api.stop_context(s.counting);
    };
    try {
        body();
    } catch (...) {
        stop();
        throw;
    }
    s.collecting = false;
    // Records arrive from the profiler's completion handler once each kernel has finished.
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::seconds(10);
    for (;;) {
        {
            std::lock_guard<std::mutex> lock(s.mu);
            if (s.completed >= s.issued) break;
        }
        if (std::chrono::steady_clock::now() > deadline) {
            stop();
            throw std::runtime_error("counter records did not arrive");
        }
        std::this_thread::sleep_for(std::chrono::milliseconds(1));
    }
    stop();
    std::lock_guard<std::mutex> lock(s.mu);
    if (missing) *missing = s.missing;
    std::vector<Dispatch> out;
    for (auto& kv : s.dispatches) out.push_back(kv.second);
    return out;
}

}  // namespace brp_counters
