// brp_workunit -- workunit driver over the HIP backend with the reference's command line,
// per-template metrics JSONL, results file, summary JSON and exit codes
// (src/workunit_main.cpp:26-152): DD observation -> device preprocessing (decode, whitening,
// zap list, optional RFI replacement) -> bank search -> toplist -> results file.
//
// Options: the reference's (required --input --bank --zap --padding --f0 --false-alarm
// --window --output --metrics; optional --wg --batch --tile --preprocessed --dump-prefix
// --reference-dir --counter-output; switches --packed --scatter --radix8 --dit --fused-power
// --radix8-store --validate --cpu), plus
//   --engine fused|plain   the batched four-step engine (default) or the reference's
//                          launch sequence
//   --group G              fused engine: templates per launch
//   --rfi-ratio R          replace bins above R x the running median (default 0: off)
// Differences, by necessity:
//   * --cpu is refused: this build's CPU path is test infrastructure, not a product mode;
//   * --validate checks each template against the other engine on the device (the
//     reference checks against its CPU path): bit-equal resampled samples, spectrum / power /
//     harmonic rel-L2 and candidates; hs_isolated_l2 is -1 and harmonic_cross_l2 holds the
//     engines' harmonic rel-L2;
//   * --batch defaults to 64 on the fused engine (1 on the plain engine, as the reference);
//     batches freeze their thresholds and replay retain() in template order, which gives the
//     one-template-at-a-time toplist;
//   * without per-template checks, the per-template wall / CPU / stage times are bank
//     averages (the fused engine times launch groups, not templates).
// Exit status 0 on success, 1 on any error.
#include <fstream>

#include "driver_kit.hpp"
#include "workload_io.hpp"

namespace fs = std::filesystem;

namespace {

struct Settings {
    std::string input, bank, zap, output, metrics;
    double padding = 0, f0 = 0, false_alarm = 0;
    uint32_t window = 0, wg = 256, tile = 0, group = 0;
    std::optional<std::string> batch, reference_dir, dump_prefix, preprocessed, counter_output;
    float rfi_ratio = 0;
    std::string engine = "fused";
    bool cpu = false, packed = false, validate = false, scatter = false, radix8 = false, stockham = true,
         fused_power = false, radix8_store = false;
};

driver::Options command_line(Settings& s) {
    driver::Options o({"option needs a value", "unknown option ", std::string("duplicate option")});
    using driver::to;
    o.required("--input", to(s.input))
        .required("--bank", to(s.bank))
        .required("--zap", to(s.zap))
        .required("--padding", to(s.padding))
        .required("--f0", to(s.f0))
        .required("--false-alarm", to(s.false_alarm))
        .required("--window", to(s.window))
        .required("--output", to(s.output))
        .required("--metrics", to(s.metrics))
        .value("--wg", to(s.wg))
        .value("--reference-dir", to(s.reference_dir))
        .value("--dump-prefix", to(s.dump_prefix))
        .value("--preprocessed", to(s.preprocessed))
        .value("--batch", to(s.batch))
        .value("--counter-output", to(s.counter_output))
        .value("--tile", to(s.tile))
        .value("--engine", to(s.engine))
        .value("--group", to(s.group))
        .value("--rfi-ratio", to(s.rfi_ratio))
        .flag("--cpu", [&] { s.cpu = true; })
        .flag("--packed", [&] { s.packed = true; })
        .flag("--validate", [&] { s.validate = true; })
        .flag("--scatter", [&] { s.scatter = true; })
        .flag("--radix8", [&] { s.radix8 = true; })
        .flag("--dit", [&] { s.stockham = false; })
        .flag("--fused-power", [&] { s.fused_power = true; })
        .flag("--radix8-store", [&] { s.radix8 = s.radix8_store = true; });
    return o;
}

// What one template's metrics line reports.
struct TemplateReport {
    size_t index = 0;
    brp::Params p{};
    double period = 0;
    bool valid = true;
    double wall_ms = 0, cpu_ms = 0;
    std::array<double, 4> stage{};
    driver::Check check{};
    double authentic_resample = -1, authentic_power = -1;
    size_t candidates = 0;
};

class MetricsFile {
public:
    MetricsFile(const fs::path& path, const Settings& s, uint32_t batch) : file_(path), s_(s), batch_(batch) {}
    void write(const TemplateReport& r) {
        const driver::Check& c = r.check;
        driver::Json j(12);
        j.put("template", r.index).put("n", r.p.n).put("steps", r.p.steps).put("wg", s_.wg).put("batch", batch_);
        j.put("packed", s_.packed).put("cpu", false).put("scatter", s_.scatter).put("radix8", s_.radix8);
        j.put("stockham", s_.stockham).put("tile", s_.tile).put("fused_power", s_.fused_power);
        j.put("radix8_store", s_.radix8_store).put("engine", s_.engine);
        j.put("period", r.period).put("tau", r.p.tau).put("psi", r.p.psi).put("s0", r.p.s0);
        j.put("validated", s_.validate).put("valid", r.valid).put("wall_ms", r.wall_ms).put("process_cpu_ms", r.cpu_ms);
        j.list("gpu_stage_ms", r.stage);
        j.put("resample_l2", c.resample_l2).put("fft_l2", c.fft_l2).put("power_l2", c.power_l2);
        j.put("hs_isolated_l2", -1).put("harmonic_cross_l2", c.harmonic_l2);
        j.put("sampled_mismatches", c.index_mismatches).put("candidate_mismatches", c.candidate_mismatches);
        j.put("authentic_resample_l2", r.authentic_resample).put("authentic_power_l2", r.authentic_power);
        j.put("candidates", r.candidates).list("thresholds", r.p.threshold);
        file_ << j.text(false) << '\n';
        file_.flush();
        if (!file_) throw std::runtime_error("metrics write failed");
    }

private:
    std::ofstream file_;
    const Settings& s_;
    uint32_t batch_;
};

void progress(size_t index, size_t total) {
    if (index % 100 == 0) std::cerr << "Processed " << index + 1 << "/" << total << " orbital templates\n";
}

int workunit(int argc, char** argv) {
    Settings s;
    command_line(s).parse(argc, argv);
    if (fs::exists(s.output) || fs::exists(s.metrics)) throw std::runtime_error("refusing to overwrite results");
    if (s.counter_output && (s.cpu || fs::exists(*s.counter_output)))
        throw std::runtime_error("counter output requires a GPU and a new file");
    if (s.cpu) throw std::runtime_error("--cpu: the CPU path of this backend is test infrastructure, not a product mode");
    if (s.engine != "fused" && s.engine != "plain") throw std::runtime_error("--engine must be fused or plain");
    const bool fused = s.engine == "fused";
    const uint32_t batch = s.batch ? uint32_t(std::stoul(*s.batch)) : (fused ? 64u : 1u);
    if ((s.radix8 && !s.stockham) || (s.tile && s.stockham))
        throw std::runtime_error("radix 8 requires Stockham; shared FFT tiles require --dit");
    const bool per_template = s.validate || s.reference_dir.has_value();
    if (!batch || batch > (fused ? 4096u : 64u) || (batch > 1 && per_template))
        throw std::runtime_error("batch must be 1..64 (fused engine: 1..4096); intermediate validation requires batch 1");

    // The observation: decoded and preprocessed on the device, or a given preprocessed series.
    const auto t_start = std::chrono::steady_clock::now();
    brp::Observation o = brp::load_observation(s.input, s.padding);
    uint32_t rfi_bins = 0;
    if (s.preprocessed) {
        o.decoded = brp::decode(o);
        o.preprocessed = brp::read_floats(*s.preprocessed, o.samples);
    } else {
        const std::vector<double> zap = brp::load_zap(s.zap);
        o.decoded.resize(o.samples);
        o.preprocessed.resize(o.samples);
        brp::hip_detail::check(brp_hip_preprocess_zap(o.block.data(), o.samples, o.bits, o.scale, o.n, s.window,
                                                      s.rfi_ratio, zap.data(), zap.size() / 2, o.sample_us,
                                                      o.padding, o.preprocessed.data(), o.decoded.data(), &rfi_bins));
    }
    const std::vector<brp::Orbit> bank = brp::load_bank(s.bank);
    const double prep_ms = driver::ms_since(t_start);

    double decoded_l2 = -1, preprocessed_l2 = -1;
    if (s.reference_dir) {
        const fs::path dir = *s.reference_dir;
        decoded_l2 = brp::relative_l2(o.decoded, brp::read_floats(dir / "decoded.f32", o.samples));
        preprocessed_l2 = brp::relative_l2(o.preprocessed, brp::read_floats(dir / "preprocessed.f32", o.samples));
        if (decoded_l2 != 0 || preprocessed_l2 >= 1e-4)
            throw std::runtime_error("authentic decoding/preprocessing validation failed");
    }
    if (s.dump_prefix) {
        brp::dump(*s.dump_prefix + "-decoded.f32", o.decoded);
        brp::dump(*s.dump_prefix + "-preprocessed.f32", o.preprocessed);
    }

    auto params = [&](const brp::Orbit& orbit) { return brp::search_parameters(o, orbit, s.f0, s.false_alarm, s.window); };
    brp_options opt{};
    opt.workgroup = s.wg;
    opt.tile = s.tile;
    opt.batch = batch;
    opt.stockham = s.stockham;
    opt.radix4 = 1;
    opt.packed = s.packed;
    opt.scatter = s.scatter;
    opt.radix8 = s.radix8;
    opt.fused_power = s.fused_power;
    opt.radix8_store = s.radix8_store;
    opt.engine = fused ? BRP_ENGINE_FUSED : BRP_ENGINE_PLAIN;
    opt.group = s.group;
    opt.graph = fused;
    opt.snapshot = per_template;
    const brp::Params first = params(bank[0]);
    brp::GPU gpu(first, o.preprocessed, opt);
    std::unique_ptr<brp::GPU> other;  // the engine --validate checks against
    if (s.validate) {
        brp_options alt = opt;
        alt.engine = fused ? BRP_ENGINE_PLAIN : BRP_ENGINE_FUSED;
        alt.graph = 0;
        alt.snapshot = 1;
        alt.tile = 0;
        alt.stockham = 1;
        other = std::make_unique<brp::GPU>(first, o.preprocessed, alt);
    }

    MetricsFile metrics(s.metrics, s, batch);
    brp::Toplist top;
    bool all_valid = true;
    double total_gpu_ms = 0;
    const auto t_bank = std::chrono::steady_clock::now();
    if (per_template) {
        // One template at a time (src/workunit_main.cpp:77-136 with batch 1), each checked.
        for (size_t index = 0; index < bank.size(); ++index) {
            TemplateReport r;
            r.index = index;
            r.period = bank[index].period;
            r.p = params(bank[index]);
            top.thresholds(r.p);
            gpu.set_template(r.p);
            const brp::Timings t = gpu.run();
            const brp::Snapshot got = gpu.snapshot();
            if (other) {
                other->set_template(r.p);
                other->run();
                r.check = driver::compare(got, other->snapshot(), r.p.steps);
            }
            if (s.reference_dir) {
                const fs::path dir = *s.reference_dir;
                const fs::path resampled = dir / ("resampled-" + std::to_string(index) + ".f32");
                if (fs::exists(resampled)) {
                    r.authentic_resample = brp::relative_l2(got.resampled, brp::read_floats(resampled, r.p.n));
                    std::vector<float> power =
                        brp::read_floats(dir / ("power-" + std::to_string(index) + ".f32"), r.p.n / 2 + 1);
                    power.resize(r.p.n / 2);  // the capture holds the Nyquist bin too
                    r.authentic_power = brp::relative_l2(got.power, power);
                }
            }
            r.valid = r.check.within_bars(false) && r.authentic_resample < 1e-4 && r.authentic_power < 1e-4;
            all_valid = all_valid && r.valid;
            top.retain(got.candidates, bank[index], r.p);
            r.wall_ms = t.wall_ms;
            r.cpu_ms = t.cpu_ms;
            r.stage = t.stage_ms;
            for (double d : r.stage) total_gpu_ms += d;
            r.candidates = got.candidates.size();
            metrics.write(r);
            progress(index, bank.size());
            if (!r.valid) throw std::runtime_error("real template correctness failure at " + std::to_string(index));
        }
    } else {
        // The bank in batches (brp_hip_bank_search): thresholds frozen per batch, retain()
        // replayed in template order, so the toplist is the one-template-at-a-time one.
        std::vector<brp::Params> ran(bank.size());
        std::vector<uint32_t> counts(bank.size());
        std::vector<brp_orbit> orbits(bank.size());
        for (size_t i = 0; i < bank.size(); ++i) orbits[i] = {bank[i].period, bank[i].tau, bank[i].psi};
        double km[4], gpu_ms = 0;
        uint64_t launches = 0, timed = 0;
        const double cpu0 = driver::usage().cpu_s;
        brp::hip_detail::check(brp_hip_kernel_ms(gpu.handle(), km, &launches, &timed));  // resets the totals
        brp::hip_detail::check(brp_hip_bank_search(gpu.handle(), orbits.data(), orbits.size(), o.samples, o.sample_us,
                                                   s.f0, s.false_alarm, s.window, top.handle(),
                                                   reinterpret_cast<brp_params*>(ran.data()), counts.data(), &gpu_ms));
        brp::hip_detail::check(brp_hip_kernel_ms(gpu.handle(), km, &launches, &timed));
        const double per = 1.0 / double(bank.size());
        const double wall_ms = driver::ms_since(t_bank) * per, cpu_ms = (driver::usage().cpu_s - cpu0) * 1e3 * per;
        std::array<double, 4> stage{};
        for (size_t k = 0; k < 4; ++k) stage[k] = timed ? km[k] / double(timed) : 0.0;
        for (double d : stage) total_gpu_ms += d * double(bank.size());
        for (size_t index = 0; index < bank.size(); ++index) {
            TemplateReport r;
            r.index = index;
            r.period = bank[index].period;
            r.p = ran[index];
            r.wall_ms = wall_ms;
            r.cpu_ms = cpu_ms;
            r.stage = stage;
            r.candidates = counts[index];
            metrics.write(r);
            progress(index, bank.size());
        }
    }
    const std::vector<brp::BankCandidate> rows = top.final();
    brp::write_candidates(s.output, rows, double(o.n) * o.sample_us * 1e-6);
    const double bank_ms = driver::ms_since(t_bank);
    if (s.counter_output) {
        std::ofstream counters(*s.counter_output);
        counters << gpu.profile_counters() << '\n';
        if (!counters) throw std::runtime_error("counter output write failed");
    }
    const driver::Usage u = driver::usage();
    driver::Json j(12);
    j.put("templates", bank.size()).put("candidates", rows.size()).put("n", o.n).put("engine", s.engine);
    j.put("preprocessing_ms", prep_ms).put("rfi_bins", rfi_bins).put("decoded_l2", decoded_l2);
    j.put("preprocessed_l2", preprocessed_l2).put("valid", all_valid).put("bank_wall_ms", bank_ms);
    j.put("total_gpu_ms", total_gpu_ms);
    j.put("templates_per_gpu_second", total_gpu_ms > 0 ? double(bank.size()) * 1000 / total_gpu_ms : 0.0);
    j.put("templates_per_wall_second", double(bank.size()) * 1000 / bank_ms);
    j.put("allocated_gpu_bytes", gpu.allocated_bytes()).put("peak_rss_kib", u.peak_rss_kib).put("process_cpu_s", u.cpu_s);
    std::cout << j.text(false) << '\n';
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    return driver::guarded("brp_workunit", [&] { return workunit(argc, argv); });
}
