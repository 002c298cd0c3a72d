"""The C++ boundary as a drop-in (CPU).

include/brp/hip.hpp replaces the reference's include/brp/vulkan.hpp.  Where the reference's
sources are present (/root/reference in the build container), its own drivers
src/main.cpp and src/workunit_main.cpp must compile against this backend with nothing changed
but the include line: g++ -std=c++20 -fsyntax-only, the reference's include directory first
(its science.hpp and workload.hpp stay the reference's own), then this repository's.  The
sources are read from /root/reference and piped to the compiler; nothing is copied.
Linking needs the reference's science and workload libraries (FFTW3f, GSL, the historical
sources), which this image lacks, so the check stops at compilation.
"""
import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
REFERENCE = pathlib.Path("/root/reference")


@pytest.mark.parametrize("driver", ["src/main.cpp", "src/workunit_main.cpp"])
def test_reference_driver_compiles_with_only_the_include_swapped(driver):
    src = REFERENCE / driver
    if not src.exists():
        pytest.skip("the reference's sources are not on this machine")
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    text = src.read_text()
    old = '#include "brp/vulkan.hpp"'
    assert text.count(old) == 1, driver
    swapped = text.replace(old, '#include "brp/hip.hpp"')
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", f"-I{REFERENCE / 'include'}", f"-I{ROOT / 'include'}",
                        "-x", "c++", "-"], input=swapped, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def test_hip_header_compiles_against_its_own_science_mirror():
    """Without the reference's headers, brp/hip.hpp takes this library's brp/hip_science.hpp
    (Params, Candidate, Orbit, Toplist, search_parameters, ...) and a reference-style bank loop
    compiles against it."""
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    code = r'''
#include "brp/hip.hpp"
#include <vector>
int main() {
    brp::Observation o;
    o.n = 1 << 16; o.samples = 1 << 16; o.sample_us = 64;
    std::vector<brp::Orbit> bank{{3600, 0.05, 1.0}};
    brp::Toplist top;
    brp::Params p = brp::search_parameters(o, bank[0], 400, 0.08, 1000);
    top.thresholds(p);
    std::vector<float> x(o.samples);
    brp::GPU gpu(p, x, 256, 0, 1, false, true, true, true);
    gpu.set_template(p);
    brp::Timings t = gpu.run();
    top.retain(gpu.candidates(0), bank[0], p);
    std::vector<brp::BankCandidate> rows = top.final();
    brp::write_candidates("out.txt", rows, o.n * o.sample_us * 1e-6);
    return int(t.stage_ms.size()) - 4;
}
'''
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", f"-I{ROOT / 'include'}", "-x", "c++", "-"],
                       input=code, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
