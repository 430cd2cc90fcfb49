"""AddressSanitizer and UndefinedBehaviorSanitizer over the compiled CPU
oracle (oracle/fdtd_cpu.c, the C restatement the bench times as its CPU
baseline and the tests use for grids the NumPy oracle cannot run in
seconds).  oracle/sanitize_check.c includes that source, runs a field model
(CPML strips, a Lorentz + Kerr slab, a source, probes) and a particle beam on
one thread and on four, and requires bitwise-equal field runs, particle runs
within 1e-12 and finite values; any out-of-bounds access, use after free,
leak or undefined operation aborts the run under -fno-sanitize-recover."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


def test_cpu_oracle_under_sanitizers(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not found")
    exe = str(tmp_path / "sanitize_check")
    cmd = [gcc, "-O1", "-g", "-fopenmp", "-ffp-contract=off", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-I", ORACLE,
           os.path.join(ORACLE, "sanitize_check.c"), "-o", exe, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="4")
    out = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert out.returncode == 0 and out.stdout.strip() == "ok", (out.returncode, out.stdout, out.stderr[-3000:])
