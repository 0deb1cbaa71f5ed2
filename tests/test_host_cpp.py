"""The C++20 host surface (ugsdr_amd/host/ugsdr_hip/*.hpp): the CRTP bases and
Hip* backends compile against include/ugsdr_hip.h (CPU), and the C++ test
program passes the reference's unit KATs and the real_8 acquisition KAT on
the GPU through both the per-operator and the fused engine, and finds the
strong satellites of the packed BbpDdc recording through the fused engine."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "test_host_api.cpp")
BIN = os.path.join(ROOT, "_build", "test_host_api")
HOST = os.path.join(ROOT, "ugsdr_amd", "host")


def build_binary():
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    headers = [os.path.join(HOST, "ugsdr_hip", f) for f in os.listdir(os.path.join(HOST, "ugsdr_hip"))]
    lib = os.path.join(ROOT, "ugsdr_amd", "libugsdr_hip.so")
    deps = [SRC, lib, *headers]
    if os.path.exists(BIN) and all(os.path.getmtime(BIN) >= os.path.getmtime(d) for d in deps):
        return
    subprocess.run(["g++", "-std=c++20", "-O2", "-Wall", "-Werror", "-I" + HOST, SRC,
                    "-L" + os.path.join(ROOT, "ugsdr_amd"), "-lugsdr_hip",
                    "-Wl,-rpath," + os.path.join(ROOT, "ugsdr_amd"), "-o", BIN], check=True)


def test_host_surface_compiles():
    build_binary()
    assert os.path.exists(BIN)


REF_SRC = os.path.join(os.environ.get("UGSDR_REFERENCE", "/root/reference"), "src")

# every Hip* backend called through its CRTP base the way the reference's
# engines call their operators (fse.hpp:253-296, tracker.hpp:106-139,
# dfe.hpp:192-203)
USE_ALL_BACKENDS = """
void use_all() {
    std::vector<std::complex<float>> v(16);
    ugsdr::HipMixer::Translate(v, 1e6, 1e3, 0.0);
    auto w = ugsdr::HipMixer::Translate(std::as_const(v), 1e6, -1e3, 0.0);
    ugsdr::HipMixer mixer(1e6, 1e3, 0.0);
    mixer.Translate(v);
    ugsdr::HipTableMixer::Translate(v, 1e6, 1e3, 0.0);
    ugsdr::HipTableMixer table_mixer(1e6, 1e3, 0.0);
    table_mixer.Translate(v);
    std::vector<float> code(16, 1.0f);
    auto spec = ugsdr::HipMatchedFilter::PrepareCodeSpectrum(code);
    ugsdr::HipMatchedFilter::FilterOptimized(v, spec);
    auto filtered = ugsdr::HipMatchedFilter::FilterOptimized(std::as_const(v), spec);
    auto a = ugsdr::HipAbs::Transform(v);
    auto r = ugsdr::HipReshapeAndSum::Transform(std::as_const(a), 4);
    ugsdr::HipReshapeAndSum::Transform(v, 4);
    auto m = ugsdr::HipMaxIndex::Transform(a);
    auto s = ugsdr::HipMeanStdDev::Calculate(a);
    auto up = ugsdr::HipUpsampler::Transform(std::as_const(code), 32);
    auto rs = ugsdr::HipResampler::Transform(std::as_const(v), 4, 8);
    auto c = ugsdr::HipCorrelator::Correlate(std::span<const std::complex<float>>(v), std::span<const float>(code));
    ugsdr::HipCorrelatorBank bank({-0.25, 0.0, 0.25});
    std::vector<float> code3(48, 1.0f);
    auto tap = ugsdr::CorrelateSplit<ugsdr::HipTrackingParametersConfig>(v, code3, 16e3, 1.0, 3.5);
    (void)tap;
    (void)w; (void)filtered; (void)r; (void)m; (void)s; (void)up; (void)rs; (void)c; (void)bank;
}
int main() { return 0; }
"""


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="no reference checkout")
def test_backends_compile_against_reference_bases():
    """UGSDR_HIP_REFERENCE_BASES: the Hip* backends and the tracking surface
    derive from the reference's own CRTP bases (src/mixer/mixer.hpp,
    src/matched_filter/matched_filter.hpp, src/math/{abs,max_index,
    mean_stddev,reshape_and_sum}.hpp, src/resample/{resampler,upsampler}.hpp,
    src/correlator/correlator.hpp) and its AcquisitionResult / Sv
    (src/acquisition/acquisition_result.hpp, src/common.hpp), and every one
    instantiates through them.  Syntax-only: nothing is linked or run."""
    ref_headers = ["mixer/mixer.hpp", "matched_filter/matched_filter.hpp", "math/abs.hpp", "math/max_index.hpp",
                   "math/mean_stddev.hpp", "math/reshape_and_sum.hpp", "resample/resampler.hpp",
                   "resample/upsampler.hpp", "correlator/correlator.hpp", "acquisition/acquisition_result.hpp"]
    # the reference headers throw std::runtime_error without including <stdexcept>
    code = "#include <stdexcept>\n" + "".join(f'#include "{h}"\n' for h in ref_headers)
    code += ('#define UGSDR_HIP_REFERENCE_BASES\n#include "ugsdr_hip/backends.hpp"\n'
             '#include "ugsdr_hip/tracking.hpp"\n') + USE_ALL_BACKENDS
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-I" + REF_SRC, "-I" + HOST, "-x", "c++", "-"],
                       input=code, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


def test_backends_instantiate_on_own_bases():
    code = '#include "ugsdr_hip/backends.hpp"\n#include "ugsdr_hip/tracking.hpp"\n' + USE_ALL_BACKENDS
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-I" + HOST, "-x", "c++", "-"], input=code,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


@pytest.mark.parametrize("call", [
    "std::vector<std::complex<double>> v(8); ugsdr::HipMixer::Translate(v, 1e6, 1e3, 0.0);",
    "std::vector<std::complex<double>> v(8); ugsdr::HipTableMixer::Translate(v, 1e6, 1e3, 0.0);",
    "std::vector<std::complex<double>> v(8); auto a = ugsdr::HipAbs::Transform(v); (void)a;",
    "std::vector<double> v(8); auto m = ugsdr::HipMaxIndex::Transform(v); (void)m;",
    "std::vector<double> v(8); auto s = ugsdr::HipMeanStdDev::Calculate(v); (void)s;",
    "std::vector<std::complex<double>> v(8); ugsdr::HipResampler::Transform(v, 4, 8);",
    "std::vector<double> c(8); auto u = ugsdr::HipUpsampler::Transform(std::as_const(c), 16); (void)u;",
])
def test_double_instantiation_stops_with_message(call):
    """the reference's typed tests also run every operator over double; the
    FP32 Hip* backends refuse that at compile time with a clear message"""
    code = '#include "ugsdr_hip/backends.hpp"\nvoid f() { ' + call + ' }\n'
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-I" + HOST, "-x", "c++", "-"], input=code,
                       capture_output=True, text=True)
    assert r.returncode != 0
    assert "compute in FP32" in r.stderr, r.stderr[-3000:]


@pytest.mark.gpu
def test_host_api_on_gpu():
    print("case C++ host API (unit KATs + real_8 per-operator vs fused)", flush=True)
    build_binary()
    r = subprocess.run([BIN, os.path.join(ROOT, "tests", "golden", "real_8_first4ms.bin"),
                        os.path.join(ROOT, "tests", "golden", "bbp_ddc_gps_L1_first4ms.bin"),
                        os.path.join(ROOT, "tests", "golden", "bbp_ddc_gps_L5_first4ms.bin")], capture_output=True,
                       text=True, timeout=600)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "host API ok" in r.stdout
