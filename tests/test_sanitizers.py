"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
library's OBJ/MTL/PNG parsers and host float32 math (a host-only g++ build of
gr_assets.cpp and gr_host.cpp) and the oracle (refcpu.c), run over the asset,
scene-file, decoration and oracle-pin tests in a child process that preloads
the sanitizer runtime, then the oracle's whole frames of the small golden
scenes and config 1 against their digests. Any report fails the test."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_LIB = os.path.join(ROOT, "gorender_amd", "csrc", "_sanitize", "libgorender_host_san.so")
SAN_ORACLE = os.path.join(ROOT, "oracle", "build", "librefcpu_san.so")
TESTS = ["tests/test_assets.py", "tests/test_scene_file.py", "tests/test_decorations.py",
         "tests/test_oracle_pins.py", "tests/test_oracle_pins_frame.py"]


def test_host_code_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "gorender_amd", "csrc"), "sanitize"],
                   check=True)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True,
                          check=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, GORENDER_HIP_LIB=SAN_LIB, REFCPU_LIB=SAN_ORACLE,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-m", "not gpu", "-p",
                        "no:cacheprovider", *TESTS], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=1200)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-6000:]
    # Whole frames: every small golden scene (clipping, overlays, ties, NaN
    # points, hand-derived frames) and config 1, skipping the large configs.
    g = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        "tests/test_golden.py", "-k", "not config2 and not config3 and not "
                        "config4 and not config5"], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=1200)
    out += g.stdout + g.stderr
    assert g.returncode == 0, out[-6000:]
    for marker in ("AddressSanitizer", "runtime error:", "LeakSanitizer"):
        assert marker not in out, out[-6000:]
    assert " passed" in out, out[-2000:]
