"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
library's host-side code in csrc/api.cpp (decomposition metadata, process
grid, status handling, profiler regions, input generators) and the oracle's
C restatement (oracle/glassman.c, oracle/decomp.c), built with
-fsanitize=address,undefined and driven over many grids by
tests/sanitize/host_checks.cpp (kernel launchers stubbed; no GPU).  A
planted out-of-bounds read shows that the build really is instrumented."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROCM = "/opt/rocm"
SAN = ["-g", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
       "-fno-omit-frame-pointer"]
INC = ["-D__HIP_PLATFORM_AMD__", "-I" + ROCM + "/include", "-I" + ROOT + "/include",
       "-I" + ROOT + "/2decomp-fft_amd/csrc", "-I" + ROOT + "/oracle"]


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    assert r.returncode == 0, " ".join(cmd) + "\n" + r.stderr[-3000:]


@pytest.fixture(scope="module")
def toolchain(tmp_path_factory):
    if not (shutil.which("gcc") and shutil.which("g++")):
        pytest.skip("gcc / g++ not found")
    if not os.path.exists(ROCM + "/include/hip/hip_runtime.h"):
        pytest.skip("HIP headers not found")
    d = tmp_path_factory.mktemp("san")
    (d / "probe.c").write_text("int main(void) { return 0; }\n")
    r = subprocess.run(["gcc"] + SAN + ["probe.c", "-o", "probe"], cwd=d, capture_output=True)
    if r.returncode != 0:
        pytest.skip("the host compiler has no sanitizer runtimes")
    return d


@pytest.mark.parametrize("sanitizer,body,report", [
    # a fixed-size array indexed past its end
    ("undefined", "int a[2] = {argc, argc}; volatile int i = 2; return a[i];",
     "index 2 out of bounds"),
    # a heap block read past its end
    ("address", "int *a = malloc(2 * sizeof(int)); a[0] = a[1] = argc; volatile int i = 2;"
     " int r = a[i]; free(a); return r;", "heap-buffer-overflow"),
])
def test_planted_overflow_is_reported(toolchain, sanitizer, body, report):
    """Each sanitizer on its own catches its planted mistake (together,
    UndefinedBehaviorSanitizer's object-size check reports the heap read
    first)."""
    d = toolchain
    (d / "oob.c").write_text("#include <stdlib.h>\nint main(int argc, char **argv) {\n"
                             "  (void)argv; " + body + "\n}\n")
    _run(["gcc", "-g", "-O1", "-fsanitize=" + sanitizer, "-fno-sanitize-recover=all",
          "oob.c", "-o", "oob"], d)
    r = subprocess.run([str(d / "oob")], capture_output=True, text=True)
    assert r.returncode != 0
    assert report in r.stderr, r.stderr


def test_host_code_and_oracle_clean_under_sanitizers(toolchain):
    d = toolchain
    for c in ("glassman.c", "decomp.c"):
        _run(["gcc"] + SAN + ["-std=c11", "-fopenmp", "-c", ROOT + "/oracle/" + c], d)
    for cpp in (ROOT + "/2decomp-fft_amd/csrc/api.cpp", ROOT + "/tests/sanitize/host_checks.cpp"):
        _run(["g++"] + SAN + ["-std=c++17", "-Wno-attributes"] + INC + ["-c", cpp], d)
    _run(["g++"] + SAN + ["-fopenmp", "host_checks.o", "api.o", "glassman.o", "decomp.o",
                          "-L" + ROCM + "/lib", "-lamdhip64", "-lrocprofiler-sdk-roctx",
                          "-Wl,-rpath," + ROCM + "/lib", "-lm", "-o", "host_checks"], d)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    r = subprocess.run([str(d / "host_checks")], capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "host checks passed" in r.stdout
    assert "runtime error" not in r.stderr
