"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

Builds tests/sanitize/host_checks.cpp with the scenario readers, the C-ABI
host runtime (its validation; no device call is made) and the CPU oracle,
all instrumented with -fsanitize=address,undefined, and runs it on the
reference's scenario files and on malformed inputs.  A sanitizer report
makes the run fail.  The HIP kernels are linked in uninstrumented (their
object from the in-tree build); nothing is launched.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "position-based-fluids_amd")
CSRC = os.path.join(PKG, "csrc")
KERNELS_OBJ = os.path.join(PKG, "_build", "pbf_kernels.o")
STATIC_OBJ = os.path.join(PKG, "_build", "pbf_static.o")
ROCM = "/opt/rocm"


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
@pytest.mark.skipif(not os.path.exists(KERNELS_OBJ),
                    reason="in-tree build missing (run __graft_entry__.build())")
def test_host_code_under_asan_and_ubsan(tmp_path, golden_dir):
    exe = str(tmp_path / "host_checks")
    flags = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
             "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
             "-fopenmp", "-D__HIP_PLATFORM_AMD__",
             "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROCM, "include")]
    srcs = [os.path.join(HERE, "sanitize", "host_checks.cpp"),
            os.path.join(CSRC, "pbf_api.cpp"),
            os.path.join(CSRC, "io_capi.cpp"),
            os.path.join(CSRC, "io", "ConfigReader.cpp"),
            os.path.join(CSRC, "io", "PartReader.cpp"),
            os.path.join(ROOT, "oracle", "pbf_oracle.cpp")]
    objs = []
    for src in srcs:
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        extra = ["-march=x86-64", "-mno-fma", "-ffp-contract=off"] \
            if src.endswith("pbf_oracle.cpp") else []
        subprocess.run(["g++", *flags, *extra, "-c", src, "-o", obj],
                       check=True, timeout=600)
        objs.append(obj)
    subprocess.run(["g++", "-fsanitize=address,undefined", "-fopenmp", "-o", exe,
                    *objs, KERNELS_OBJ, STATIC_OBJ, "-L" + os.path.join(ROCM, "lib"),
                    "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib")],
                   check=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, golden_dir, str(tmp_path)], env=env,
                       capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert "0 failure(s)" in out
