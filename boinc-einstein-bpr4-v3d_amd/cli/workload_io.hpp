// workload_io.hpp -- the workunit's file formats (src/workload.cpp:29-56 and 96-107,
// src/workunit_main.cpp:15-25): the DD observation (gzip or plain, a 1168-byte header of
// little-endian doubles and the sample count, then 4-bit or 8-bit samples), the orbital
// bank (one "period tau psi" row per line), the zap list ("lower upper" Hz pairs) and raw
// float32 files.  Host code of the command-line drivers over brp/hip_science.hpp's
// Observation; the decode and preprocessing run on the device (brp_hip_preprocess_zap).
#pragma once
#include <zlib.h>

#include <algorithm>
#include <array>
#include <bit>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "brp/hip.hpp"
#include "brp/hip_science.hpp"

namespace brp {

// src/workload.cpp:29-56: header fields sample_us (double 0), observation_s (double 1),
// dm (double 13), scale (double 14), samples (uint32 at byte 128); n = padding *
// samples rounded; .bin4 packs two 4-bit samples per byte (high nibble first), .binary
// holds signed bytes; no byte may follow the samples.
inline Observation load_observation(const std::filesystem::path& path, double padding) {
    static_assert(std::endian::native == std::endian::little, "DD loader currently requires little-endian host");
    gzFile stream = gzopen(path.c_str(), "rb");
    if (!stream) throw std::runtime_error("cannot open DD input");
    struct Close {
        gzFile s;
        ~Close() { gzclose(s); }
    } close{stream};
    std::array<unsigned char, 1168> header{};
    if (gzread(stream, header.data(), unsigned(header.size())) != int(header.size()))
        throw std::runtime_error("truncated DD header");
    auto field = [&](size_t i) {
        double v;
        std::memcpy(&v, header.data() + 8 * i, 8);
        return v;
    };
    Observation o;
    o.sample_us = field(0);
    o.observation_s = field(1);
    o.dm = field(13);
    o.scale = field(14);
    o.padding = padding;
    std::memcpy(&o.samples, header.data() + 128, 4);
    if (!o.samples || o.samples > (1u << 24) || !std::isfinite(o.scale) || o.scale <= 0 ||
        !std::isfinite(o.sample_us) || o.sample_us <= 0 || !std::isfinite(padding) || padding < 1 || padding > 10)
        throw std::runtime_error("invalid DD metadata");
    const double size = padding * o.samples + 0.5;
    if (size > 3 * (1u << 22) + 0.5) throw std::runtime_error("observation exceeds supported length");
    o.n = uint32_t(size);
    const bool packed = path.extension() == ".bin4";
    if (!packed && path.extension() != ".binary") throw std::runtime_error("unsupported DD sample format");
    if (packed && (o.samples & 1)) throw std::runtime_error("odd packed DD length");
    o.bits = packed ? 4 : 8;
    o.block.resize(packed ? o.samples / 2 : o.samples);
    if (gzread(stream, o.block.data(), unsigned(o.block.size())) != int(o.block.size()))
        throw std::runtime_error("truncated DD samples");
    unsigned char extra;
    if (gzread(stream, &extra, 1) != 0) throw std::runtime_error("trailing DD bytes or gzip error");
    return o;
}

// The decode of src/workload.cpp:50-54 on the host (the device preprocessing decodes
// the same way): float(double(v) / scale), 4-bit high nibble first or signed bytes.
inline std::vector<float> decode(const Observation& o) {
    std::vector<float> out(o.samples);
    for (uint32_t i = 0; i < o.samples; ++i) {
        const double v = o.bits == 4 ? double((i & 1) ? (o.block[i / 2] & 15) : (o.block[i / 2] >> 4))
                                     : double(int8_t(o.block[i]));
        out[i] = float(v / o.scale);
    }
    return out;
}

// src/workload.cpp:96-107
inline std::vector<Orbit> load_bank(const std::filesystem::path& path) {
    std::ifstream file(path);
    if (!file) throw std::runtime_error("cannot open orbital bank");
    std::vector<Orbit> out;
    std::string line;
    while (std::getline(file, line)) {
        Orbit o{};
        std::string extra;
        std::istringstream row(line);
        if (!(row >> o.period >> o.tau >> o.psi) || (row >> extra) || !std::isfinite(o.period + o.tau + o.psi) ||
            o.period <= 0 || o.tau < 0)
            throw std::runtime_error("invalid orbital bank row");
        out.push_back(o);
    }
    if (out.empty()) throw std::runtime_error("empty orbital bank");
    return out;
}

// The zap list of src/workload.cpp:76-85: whitespace-separated (lower, upper) pairs in Hz,
// flattened; the ranges themselves are checked by the preprocessing.
inline std::vector<double> load_zap(const std::filesystem::path& path) {
    std::ifstream file(path);
    if (!file) throw std::runtime_error("cannot open zap list");
    std::vector<double> out;
    double lower, upper;
    while (file >> lower >> upper) {
        out.push_back(lower);
        out.push_back(upper);
    }
    if (!file.eof()) throw std::runtime_error("malformed zap list");
    return out;
}

// src/workunit_main.cpp:15-21
inline std::vector<float> read_floats(const std::filesystem::path& path, size_t count) {
    std::ifstream file(path, std::ios::binary | std::ios::ate);
    if (!file || size_t(file.tellg()) != count * 4) throw std::runtime_error("reference float dimensions: " + path.string());
    std::vector<float> values(count);
    file.seekg(0);
    file.read(reinterpret_cast<char*>(values.data()), std::streamsize(count * 4));
    if (!file) throw std::runtime_error("reference read failed");
    return values;
}

// src/workunit_main.cpp:22-25
inline void dump(const std::filesystem::path& path, std::span<const float> values) {
    std::ofstream file(path, std::ios::binary);
    file.write(reinterpret_cast<const char*>(values.data()), std::streamsize(values.size_bytes()));
    if (!file) throw std::runtime_error("state dump failed");
}

}  // namespace brp
