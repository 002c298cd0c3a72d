// driver_kit.hpp -- the pieces both command-line drivers share: a declarative option table,
// a small JSON writer, the engine-against-engine result check and timing statistics.
//
// The drivers keep the reference's command lines, JSON field names, messages and exit codes
// (src/main.cpp:13-37, 96-118; src/workunit_main.cpp:27-52, 125-150), so scripts written for
// brp_bench and brp_workunit keep working; the code that provides them is this file's.
#pragma once
#include <sys/resource.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <functional>
#include <iostream>
#include <map>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "brp/hip.hpp"

namespace driver {

// ------------------------------------------------------------------ options --
// A command line is a list of switches (no value) and valued options.  Each program states
// its own table and its parse policy: the messages for a missing value and an unknown name,
// whether a repeated option is an error or overrides, and which options are required.
struct Policy {
    std::string missing_value;        // a valued option ends the command line
    std::string unknown_prefix;       // + the option's name
    std::optional<std::string> twice; // a repeated valued option (error message), else the last wins
};

class Options {
public:
    using Apply = std::function<void(const std::string&)>;

    explicit Options(Policy policy) : policy_(std::move(policy)) {}

    Options& flag(std::string name, std::function<void()> apply, std::string usage = {}) {
        items_.push_back({std::move(name), false, false, [a = std::move(apply)](const std::string&) { a(); },
                          std::move(usage)});
        return *this;
    }
    Options& value(std::string name, Apply apply, std::string usage = {}, bool required = false) {
        items_.push_back({std::move(name), true, required, std::move(apply), std::move(usage)});
        return *this;
    }
    Options& required(std::string name, Apply apply) { return value(std::move(name), std::move(apply), {}, true); }

    // "program [usage] [usage] ..." in table order, for the options that have a usage text.
    std::string usage(const std::string& program) const {
        std::string line = program;
        for (const auto& it : items_)
            if (!it.usage.empty()) line += " [" + it.usage + "]";
        return line;
    }

    // Reads argv: switches apply at once; values are collected, then the required ones are
    // checked (in table order), then unknown names, then every value applies in table order.
    // Returns the names given (switches and values).
    std::vector<std::string> parse(int argc, char** argv) {
        std::map<std::string, std::string> given;
        std::vector<std::string> order, names;
        for (int i = 1; i < argc; ++i) {
            const std::string name = argv[i];
            names.push_back(name);
            if (const Item* it = find(name); it && !it->valued) {
                it->apply({});
                continue;
            }
            if (i + 1 >= argc) throw std::runtime_error(policy_.missing_value);
            const std::string v = argv[++i];
            if (given.count(name) && policy_.twice) throw std::runtime_error(*policy_.twice);
            if (!given.count(name)) order.push_back(name);
            given[name] = v;
        }
        for (const auto& it : items_)
            if (it.required && !given.count(it.name)) throw std::runtime_error("required option " + it.name);
        for (const auto& name : order)
            if (!find(name)) throw std::runtime_error(policy_.unknown_prefix + name);
        for (const auto& it : items_)
            if (it.valued && given.count(it.name)) it.apply(given.at(it.name));
        return names;
    }

private:
    struct Item {
        std::string name;
        bool valued, required;
        Apply apply;
        std::string usage;
    };
    const Item* find(const std::string& name) const {
        for (const auto& it : items_)
            if (it.name == name) return &it;
        return nullptr;
    }
    Policy policy_;
    std::vector<Item> items_;
};

// Setters for numeric options, with the standard library's parse errors.
inline Options::Apply to(uint32_t& dst) { return [&dst](const std::string& v) { dst = uint32_t(std::stoul(v)); }; }
inline Options::Apply to(float& dst) { return [&dst](const std::string& v) { dst = std::stof(v); }; }
inline Options::Apply to(double& dst) { return [&dst](const std::string& v) { dst = std::stod(v); }; }
inline Options::Apply to(std::string& dst) { return [&dst](const std::string& v) { dst = v; }; }
inline Options::Apply to(std::optional<std::string>& dst) {
    return [&dst](const std::string& v) { dst = v; };
}

// --------------------------------------------------------------------- JSON --
// An object whose members are written in insertion order.  Numbers use a fixed significant
// digit count; a value that is not finite becomes null (JSON has no inf / nan).
class Json {
public:
    explicit Json(int digits = 10) : digits_(digits) {}

    Json& put(const std::string& key, double v) { return raw(key, number(v)); }
    Json& put(const std::string& key, float v) { return put(key, double(v)); }
    Json& put(const std::string& key, bool v) { return raw(key, v ? "true" : "false"); }
    Json& put(const std::string& key, const char* v) { return raw(key, quote(v)); }
    Json& put(const std::string& key, const std::string& v) { return raw(key, quote(v)); }
    Json& put(const std::string& key, std::nullptr_t) { return raw(key, "null"); }
    Json& put(const std::string& key, const Json& v) { return raw(key, v.text(false)); }
    template <class T>
    std::enable_if_t<std::is_integral_v<T> && !std::is_same_v<T, bool>, Json&> put(const std::string& key, T v) {
        return raw(key, std::to_string(v));
    }
    template <class T>
    Json& list(const std::string& key, const T& values) {
        std::string s = "[";
        for (const auto& v : values) s += (s.size() > 1 ? "," : "") + number(double(v));
        return raw(key, s + "]");
    }
    // A member whose value is already JSON text (for example the counter report).
    Json& raw(const std::string& key, std::string json_text) {
        members_.emplace_back(key, std::move(json_text));
        return *this;
    }
    // One line, or one member per line.
    std::string text(bool multiline) const {
        std::string s = "{";
        for (size_t i = 0; i < members_.size(); ++i) {
            s += i ? "," : "";
            if (multiline) s += "\n  ";
            s += quote(members_[i].first) + (multiline ? ": " : ":") + members_[i].second;
        }
        return s + (multiline ? "\n}" : "}");
    }

private:
    std::string number(double v) const {
        if (!std::isfinite(v)) return "null";
        std::ostringstream o;
        o.precision(digits_);
        o << v;
        return o.str();
    }
    static std::string quote(const std::string& s) {
        std::string q = "\"";
        for (char c : s) {
            if (c == '"' || c == '\\') q += '\\';
            if (uint8_t(c) < 0x20) {
                char buf[8];
                std::snprintf(buf, sizeof buf, "\\u%04x", c);
                q += buf;
                continue;
            }
            q += c;
        }
        return q + "\"";
    }
    int digits_;
    std::vector<std::pair<std::string, std::string>> members_;
};

// ----------------------------------------------------------- result check --
// One template's result against a second computation of it (here the other engine on the
// device), with the reference's bars (src/main.cpp:55-80): bit-equal resampled samples,
// rel-L2 of the series, spectrum, power and harmonic sums, the candidate lists (same bins and
// levels, power within 2e-4 (1 + |p|)) and the top 100 of each level ranked by power.
struct Check {
    uint32_t index_mismatches = 0, candidate_mismatches = 0, rank_mismatches = 0;
    double resample_l2 = -1, fft_l2 = -1, power_l2 = -1, harmonic_l2 = -1;

    bool within_bars(bool ranks) const {
        return !index_mismatches && !candidate_mismatches && (!ranks || !rank_mismatches) && resample_l2 < 5e-5 &&
               fft_l2 < 3e-5 && power_l2 < 5e-5 && harmonic_l2 < 5e-5;
    }
};

inline std::span<const float> as_floats(const std::vector<std::complex<float>>& v) {
    return {reinterpret_cast<const float*>(v.data()), 2 * v.size()};
}

// Bins of a level's candidates, strongest first (equal powers: lower bin first).
inline std::vector<uint32_t> ranked_bins(const std::vector<brp::Candidate>& all, uint32_t level, size_t top) {
    std::vector<std::pair<float, uint32_t>> v;
    for (const auto& c : all)
        if (c.level == level) v.emplace_back(c.power, c.bin);
    std::sort(v.begin(), v.end(), [](const auto& a, const auto& b) {
        return std::make_pair(-a.first, a.second) < std::make_pair(-b.first, b.second);
    });
    std::vector<uint32_t> bins;
    for (size_t i = 0; i < std::min(top, v.size()); ++i) bins.push_back(v[i].second);
    return bins;
}

inline Check compare(const brp::Snapshot& got, const brp::Snapshot& want, uint32_t steps) {
    Check c;
    for (uint32_t i = 0; i < steps; ++i) c.index_mismatches += got.resampled[i] != want.resampled[i];
    c.resample_l2 = brp::relative_l2(got.resampled, want.resampled);
    c.fft_l2 = brp::relative_l2(as_floats(got.fft), as_floats(want.fft));
    c.power_l2 = brp::relative_l2(got.power, want.power);
    c.harmonic_l2 = brp::relative_l2(got.summed, want.summed);
    // Both lists are in (level, bin) order: pairwise, or one mismatch for different lengths.
    if (got.candidates.size() != want.candidates.size()) {
        c.candidate_mismatches = 1;
    } else {
        for (size_t i = 0; i < got.candidates.size(); ++i) {
            const brp::Candidate& a = got.candidates[i];
            const brp::Candidate& b = want.candidates[i];
            const bool same = a.bin == b.bin && a.level == b.level &&
                              std::fabs(a.power - b.power) <= 2e-4f * (1.0f + std::fabs(b.power));
            c.candidate_mismatches += !same;
        }
    }
    for (uint32_t h = 0; h < 5; ++h) {
        const auto a = ranked_bins(got.candidates, h, 100), b = ranked_bins(want.candidates, h, 100);
        for (size_t i = 0; i < std::min(a.size(), b.size()); ++i) c.rank_mismatches += a[i] != b[i];
    }
    return c;
}

// A replay's snapshot equals an earlier one bit for bit.
inline bool identical(const brp::Snapshot& a, const brp::Snapshot& b) {
    if (a.resampled != b.resampled || a.fft != b.fft || a.power != b.power || a.summed != b.summed ||
        a.candidates.size() != b.candidates.size())
        return false;
    for (size_t i = 0; i < a.candidates.size(); ++i) {
        const auto &x = a.candidates[i], &y = b.candidates[i];
        if (x.bin != y.bin || x.level != y.level || x.power != y.power) return false;
    }
    return true;
}

// --------------------------------------------------------------- timing --
struct Series {
    std::vector<double> v;
    void add(double x) { v.push_back(x); }
    double median() const {
        std::vector<double> s = v;
        std::nth_element(s.begin(), s.begin() + s.size() / 2, s.end());
        return s[s.size() / 2];
    }
    double min() const { return *std::min_element(v.begin(), v.end()); }
    double max() const { return *std::max_element(v.begin(), v.end()); }
};

inline double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

struct Usage {
    long peak_rss_kib;
    double cpu_s;
};
inline Usage usage() {
    rusage u{};
    getrusage(RUSAGE_SELF, &u);
    return {u.ru_maxrss, double(u.ru_utime.tv_sec) + u.ru_utime.tv_usec / 1e6 + double(u.ru_stime.tv_sec) +
                             u.ru_stime.tv_usec / 1e6};
}

// Runs `body`; an exception ends the program with "<program>: <message>" and status 1 (the
// drivers' error exit, src/main.cpp:117-118).
inline int guarded(const char* program, const std::function<int()>& body) {
    try {
        return body();
    } catch (const std::exception& e) {
        std::cerr << program << ": " << e.what() << '\n';
        return 1;
    }
}

}  // namespace driver
