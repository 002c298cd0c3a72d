// brp_workunit -- the reference's workunit driver (src/workunit_main.cpp:26-152) on the
// HIP backend: DD observation -> device preprocessing (decode, whitening, zap list,
// optional RFI replacement) -> bank search -> toplist -> results file, with one metrics
// JSON object per template and a summary JSON line on stdout.
//
// Options are the reference's (required: --input --bank --zap --padding --f0
// --false-alarm --window --output --metrics; optional: --wg --batch --tile --preprocessed
// --dump-prefix --reference-dir --counter-output; switches --packed --scatter --radix8
// --dit --fused-power --radix8-store --validate --cpu), plus this backend's:
//   --engine fused|plain   fused: the batched four-step engine (default); plain: the
//                          reference's launch sequence
//   --group G              fused engine: templates per launch (default 8)
//   --rfi-ratio R          replace bins above R x the running median (default 0: off)
// Differences, by necessity:
//   * --cpu is refused: the CPU path of this build is test infrastructure only;
//   * --validate checks each template against the other engine on the device (the
//     reference checks against its CPU path): bit-equal resample, FFT / power / harmonic
//     rel-L2 and candidates; hs_isolated_l2 is reported as -1;
//   * --batch defaults to 64 with the fused engine (1 with the plain engine, as the
//     reference); thresholds are frozen per batch and retain() replayed in template
//     order, which gives the one-template-at-a-time toplist;
//   * per-template wall / CPU / stage times are bank averages (the fused engine times
//     launch groups, not templates).
// Exit status 0 on success, 1 on any error (src/workunit_main.cpp:152).
#include <sys/resource.h>

#include <chrono>
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <map>

#include "workload_io.hpp"

namespace fs = std::filesystem;

namespace {

double process_cpu_s() {
    rusage u{};
    getrusage(RUSAGE_SELF, &u);
    return double(u.ru_utime.tv_sec) + u.ru_utime.tv_usec / 1e6 + double(u.ru_stime.tv_sec) + u.ru_stime.tv_usec / 1e6;
}

long peak_rss_kib() {
    rusage u{};
    getrusage(RUSAGE_SELF, &u);
    return u.ru_maxrss;
}

const char* tf(bool b) { return b ? "true" : "false"; }

std::span<const float> floats(const std::vector<std::complex<float>>& x) {
    return {reinterpret_cast<const float*>(x.data()), 2 * x.size()};
}

}  // namespace

int main(int argc, char** argv) try {
    std::map<std::string, std::string> options;
    bool cpu = false, packed = false, validate = false, scatter = false, radix8 = false, stockham = true,
         fused_power = false, radix8_store = false;
    for (int i = 1; i < argc; ++i) {
        const std::string key = argv[i];
        if (key == "--cpu") { cpu = true; continue; }
        if (key == "--packed") { packed = true; continue; }
        if (key == "--validate") { validate = true; continue; }
        if (key == "--scatter") { scatter = true; continue; }
        if (key == "--radix8") { radix8 = true; continue; }
        if (key == "--dit") { stockham = false; continue; }
        if (key == "--fused-power") { fused_power = true; continue; }
        if (key == "--radix8-store") { radix8_store = true; radix8 = true; continue; }
        if (i + 1 == argc) throw std::runtime_error("option needs a value");
        if (!options.emplace(key, argv[++i]).second) throw std::runtime_error("duplicate option");
    }
    for (const char* key : {"--input", "--bank", "--zap", "--padding", "--f0", "--false-alarm", "--window", "--output",
                            "--metrics"})
        if (!options.contains(key)) throw std::runtime_error(std::string("required option ") + key);
    static const char* known[] = {"--input",  "--bank",          "--zap",         "--padding",      "--f0",
                                  "--false-alarm", "--window",   "--output",      "--metrics",      "--wg",
                                  "--reference-dir", "--dump-prefix", "--preprocessed", "--batch", "--counter-output",
                                  "--tile",   "--engine",        "--group",       "--rfi-ratio"};
    for (const auto& [key, value] : options)
        if (std::find_if(std::begin(known), std::end(known), [&](const char* k) { return key == k; }) == std::end(known))
            throw std::runtime_error("unknown option " + key);
    if (fs::exists(options.at("--output")) || fs::exists(options.at("--metrics")))
        throw std::runtime_error("refusing to overwrite results");
    if (options.contains("--counter-output") && (cpu || fs::exists(options.at("--counter-output"))))
        throw std::runtime_error("counter output requires a GPU and a new file");
    if (cpu) throw std::runtime_error("--cpu: the CPU path of this backend is test infrastructure, not a product mode");
    const std::string engine_name = options.contains("--engine") ? options.at("--engine") : "fused";
    if (engine_name != "fused" && engine_name != "plain") throw std::runtime_error("--engine must be fused or plain");
    const uint32_t engine = engine_name == "fused" ? BRP_ENGINE_FUSED : BRP_ENGINE_PLAIN;

    const auto start = std::chrono::steady_clock::now();
    brp::Observation o = brp::load_observation(options.at("--input"), std::stod(options.at("--padding")));
    const uint32_t window = std::stoul(options.at("--window"));
    const uint32_t wg = options.contains("--wg") ? std::stoul(options.at("--wg")) : 256;
    const uint32_t batch = options.contains("--batch") ? std::stoul(options.at("--batch"))
                                                       : (engine == BRP_ENGINE_FUSED ? 64 : 1);
    const uint32_t tile = options.contains("--tile") ? std::stoul(options.at("--tile")) : 0;
    const uint32_t group = options.contains("--group") ? std::stoul(options.at("--group")) : 0;
    const float rfi_ratio = options.contains("--rfi-ratio") ? std::stof(options.at("--rfi-ratio")) : 0.0f;
    const double f0 = std::stod(options.at("--f0")), false_alarm = std::stod(options.at("--false-alarm"));
    if ((radix8 && !stockham) || (tile && stockham))
        throw std::runtime_error("radix 8 requires Stockham; shared FFT tiles require --dit");
    if (!batch || batch > (engine == BRP_ENGINE_FUSED ? 4096u : 64u) ||
        (batch > 1 && (validate || options.contains("--reference-dir"))))
        throw std::runtime_error("batch must be 1..64 (fused engine: 1..4096); intermediate validation requires batch 1");

    // Decode + preprocessing on the device, or the given preprocessed series
    // (src/workunit_main.cpp:47-48).
    uint32_t rfi_bins = 0;
    if (options.contains("--preprocessed")) {
        o.decoded = brp::decode(o);
        o.preprocessed = brp::read_floats(options.at("--preprocessed"), o.samples);
    } else {
        const std::vector<double> zap = brp::load_zap(options.at("--zap"));
        o.preprocessed = brp::preprocess(o.block, o.samples, o.bits, o.scale, o.n, window, rfi_ratio, zap,
                                         o.sample_us, o.padding, &o.decoded, &rfi_bins);
    }
    const std::vector<brp::Orbit> bank = brp::load_bank(options.at("--bank"));
    const double prep_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();

    double decoded_error = -1, preprocessed_error = -1;
    if (options.contains("--reference-dir")) {
        const fs::path folder = options.at("--reference-dir");
        decoded_error = brp::relative_l2(o.decoded, brp::read_floats(folder / "decoded.f32", o.samples));
        preprocessed_error = brp::relative_l2(o.preprocessed, brp::read_floats(folder / "preprocessed.f32", o.samples));
        if (decoded_error != 0 || preprocessed_error >= 1e-4)
            throw std::runtime_error("authentic decoding/preprocessing validation failed");
    }
    if (options.contains("--dump-prefix")) {
        brp::dump(options.at("--dump-prefix") + "-decoded.f32", o.decoded);
        brp::dump(options.at("--dump-prefix") + "-preprocessed.f32", o.preprocessed);
    }

    auto params = [&](const brp::Orbit& orbit) {
        return brp::search_parameters(o.n, o.samples, o.sample_us, orbit, f0, false_alarm, window);
    };
    const brp::Params first = params(bank[0]);
    brp_options opt{};
    opt.workgroup = wg; opt.tile = tile; opt.batch = batch; opt.stockham = stockham; opt.radix4 = 1;
    opt.packed = packed; opt.scatter = scatter; opt.radix8 = radix8; opt.fused_power = fused_power;
    opt.radix8_store = radix8_store; opt.engine = engine; opt.group = group;
    opt.graph = engine == BRP_ENGINE_FUSED;
    const bool per_template = validate || options.contains("--reference-dir");
    opt.snapshot = per_template;
    brp::GPU gpu(first, o.preprocessed, opt);
    std::unique_ptr<brp::GPU> other;  // the cross-check engine of --validate
    if (validate) {
        brp_options oo = opt;
        oo.engine = engine == BRP_ENGINE_FUSED ? BRP_ENGINE_PLAIN : BRP_ENGINE_FUSED;
        oo.graph = 0;
        oo.snapshot = 1;
        oo.tile = 0;
        oo.stockham = 1;
        other = std::make_unique<brp::GPU>(first, o.preprocessed, oo);
    }

    std::ofstream metrics(options.at("--metrics"));
    metrics << std::setprecision(12);
    brp::Toplist top;
    bool all_valid = true;
    const auto bank_start = std::chrono::steady_clock::now();
    const double cpu_start = process_cpu_s();
    double total_gpu_ms = 0;
    auto emit = [&](size_t index, const brp::Params& p, bool valid, double wall_ms, double cpu_ms,
                    const std::array<double, 4>& stage, double resample_error, double fft_error, double power_error,
                    double hs_cross, uint32_t sampled_mismatches, uint32_t candidate_mismatches, double auth_resample,
                    double auth_power, size_t candidates) {
        metrics << "{\"template\":" << index << ",\"n\":" << p.n << ",\"steps\":" << p.steps << ",\"wg\":" << wg
                << ",\"batch\":" << batch << ",\"packed\":" << tf(packed) << ",\"cpu\":false"
                << ",\"scatter\":" << tf(scatter) << ",\"radix8\":" << tf(radix8) << ",\"stockham\":" << tf(stockham)
                << ",\"tile\":" << tile << ",\"fused_power\":" << tf(fused_power) << ",\"radix8_store\":"
                << tf(radix8_store) << ",\"engine\":\"" << engine_name << "\",\"period\":" << bank[index].period
                << ",\"tau\":" << p.tau << ",\"psi\":" << p.psi << ",\"s0\":" << p.s0 << ",\"validated\":"
                << tf(validate) << ",\"valid\":" << tf(valid) << ",\"wall_ms\":" << wall_ms
                << ",\"process_cpu_ms\":" << cpu_ms << ",\"gpu_stage_ms\":[";
        for (size_t i = 0; i < 4; ++i) metrics << (i ? "," : "") << stage[i];
        metrics << "],\"resample_l2\":" << resample_error << ",\"fft_l2\":" << fft_error << ",\"power_l2\":"
                << power_error << ",\"hs_isolated_l2\":-1,\"harmonic_cross_l2\":" << hs_cross
                << ",\"sampled_mismatches\":" << sampled_mismatches << ",\"candidate_mismatches\":"
                << candidate_mismatches << ",\"authentic_resample_l2\":" << auth_resample
                << ",\"authentic_power_l2\":" << auth_power << ",\"candidates\":" << candidates << ",\"thresholds\":[";
        for (size_t h = 0; h < 5; ++h) metrics << (h ? "," : "") << p.threshold[h];
        metrics << "]}\n";
        metrics.flush();
    };

    if (per_template) {
        // src/workunit_main.cpp:77-136 one template at a time (batch 1), each checked.
        for (size_t index = 0; index < bank.size(); ++index) {
            brp::Params p = params(bank[index]);
            top.thresholds(p);
            gpu.set_template(p);
            const brp::Timings timing = gpu.run();
            const brp::Snapshot actual = gpu.snapshot();
            double resample_error = -1, fft_error = -1, power_error = -1, hs_cross = -1;
            uint32_t sampled_mismatches = 0, candidate_mismatches = 0;
            if (other) {
                other->set_template(p);
                other->run();
                const brp::Snapshot ref = other->snapshot();
                for (uint32_t i = 0; i < p.steps; ++i) sampled_mismatches += actual.resampled[i] != ref.resampled[i];
                resample_error = brp::relative_l2(actual.resampled, ref.resampled);
                fft_error = brp::relative_l2(floats(actual.fft), floats(ref.fft));
                power_error = brp::relative_l2(actual.power, ref.power);
                hs_cross = brp::relative_l2(actual.summed, ref.summed);
                if (actual.candidates.size() != ref.candidates.size()) ++candidate_mismatches;
                else
                    for (size_t i = 0; i < actual.candidates.size(); ++i) {
                        const auto a = actual.candidates[i], b = ref.candidates[i];
                        if (a.bin != b.bin || a.level != b.level ||
                            std::abs(a.power - b.power) > 2e-4 * (1 + std::abs(b.power)))
                            ++candidate_mismatches;
                    }
            }
            double auth_resample = -1, auth_power = -1;
            if (options.contains("--reference-dir")) {
                const fs::path folder = options.at("--reference-dir");
                const fs::path rs = folder / ("resampled-" + std::to_string(index) + ".f32");
                if (fs::exists(rs)) {
                    auth_resample = brp::relative_l2(actual.resampled, brp::read_floats(rs, p.n));
                    auto power = brp::read_floats(folder / ("power-" + std::to_string(index) + ".f32"), p.n / 2 + 1);
                    power.pop_back();
                    auth_power = brp::relative_l2(actual.power, power);
                }
            }
            const bool valid = !sampled_mismatches && !candidate_mismatches && resample_error < 5e-5 &&
                               fft_error < 3e-5 && power_error < 5e-5 && hs_cross < 5e-5 && auth_resample < 1e-4 &&
                               auth_power < 1e-4;
            all_valid = all_valid && valid;
            top.retain(actual.candidates, bank[index], p);
            std::array<double, 4> stage{timing.stage_ms[0], timing.stage_ms[1], timing.stage_ms[2], timing.stage_ms[3]};
            for (double d : stage) total_gpu_ms += d;
            emit(index, p, valid, timing.wall_ms, timing.cpu_ms, stage, resample_error, fft_error, power_error, hs_cross,
                 sampled_mismatches, candidate_mismatches, auth_resample, auth_power, actual.candidates.size());
            if (index % 100 == 0) std::cerr << "Processed " << index + 1 << "/" << bank.size() << " orbital templates\n";
            if (!valid) throw std::runtime_error("real template correctness failure at " + std::to_string(index));
        }
    } else {
        // The bank loop in batches with thresholds frozen per batch and retain() in
        // template order (brp_hip_bank_search), equal to the one-at-a-time toplist.
        std::vector<brp::Params> run(bank.size());
        std::vector<uint32_t> counts(bank.size());
        double gpu_ms = 0;
        double km[4];
        uint64_t launches = 0, timed = 0;
        brp::check(brp_hip_kernel_ms(gpu.handle(), km, &launches, &timed));  // reset
        brp::check(brp_hip_bank_search(gpu.handle(), bank.data(), bank.size(), o.samples, o.sample_us, f0, false_alarm,
                                       window, top.handle(), run.data(), counts.data(), &gpu_ms));
        brp::check(brp_hip_kernel_ms(gpu.handle(), km, &launches, &timed));
        const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - bank_start).count();
        const double cpu_ms = (process_cpu_s() - cpu_start) * 1e3;
        std::array<double, 4> stage{};
        for (int s = 0; s < 4; ++s) stage[s] = timed ? km[s] / double(timed) : 0.0;
        for (double d : stage) total_gpu_ms += d * double(bank.size());
        for (size_t index = 0; index < bank.size(); ++index) {
            emit(index, run[index], true, wall / double(bank.size()), cpu_ms / double(bank.size()), stage, -1, -1, -1,
                 -1, 0, 0, -1, -1, counts[index]);
            if (index % 100 == 0) std::cerr << "Processed " << index + 1 << "/" << bank.size() << " orbital templates\n";
        }
    }
    const std::vector<brp::BankCandidate> result = top.final();
    brp::write_candidates(options.at("--output"), result, double(o.n) * o.sample_us * 1e-6);
    const double bank_wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - bank_start).count();
    if (options.contains("--counter-output")) {
        std::ofstream counters(options.at("--counter-output"));
        counters << gpu.profile_counters() << '\n';
        if (!counters) throw std::runtime_error("counter output write failed");
    }
    std::cout << std::setprecision(12) << "{\"templates\":" << bank.size() << ",\"candidates\":" << result.size()
              << ",\"n\":" << o.n << ",\"engine\":\"" << engine_name << "\",\"preprocessing_ms\":" << prep_ms
              << ",\"rfi_bins\":" << rfi_bins << ",\"decoded_l2\":" << decoded_error << ",\"preprocessed_l2\":"
              << preprocessed_error << ",\"valid\":" << tf(all_valid) << ",\"bank_wall_ms\":" << bank_wall_ms
              << ",\"total_gpu_ms\":" << total_gpu_ms << ",\"templates_per_gpu_second\":"
              << (total_gpu_ms > 0 ? double(bank.size()) * 1000 / total_gpu_ms : 0)
              << ",\"templates_per_wall_second\":" << double(bank.size()) * 1000 / bank_wall_ms
              << ",\"allocated_gpu_bytes\":" << gpu.allocated_bytes() << ",\"peak_rss_kib\":" << peak_rss_kib()
              << ",\"process_cpu_s\":" << process_cpu_s() << "}\n";
    return 0;
} catch (const std::exception& e) {
    std::cerr << "brp_workunit: " << e.what() << '\n';
    return 1;
}
