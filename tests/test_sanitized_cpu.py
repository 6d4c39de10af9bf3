"""CPU tests under sanitizers. The server's upload staging bookkeeping
(page-locked blocks held until their copies complete;
tgmm_amd/csrc/tgmm_staging.h) runs with plain memory and fake events
(tests/native/staging_test.cpp), built plain and with AddressSanitizer and
UndefinedBehaviorSanitizer; the oracle runs under both sanitizers on small
frames (tests/native/oracle_san.c)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "staging_test.cpp")


@pytest.mark.parametrize("san", [[], ["-fsanitize=address,undefined", "-fno-omit-frame-pointer"]])
def test_staging_bookkeeping(tmp_path, san):
    exe = str(tmp_path / "staging_test")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g"] + san + ["-o", exe, SRC],
                       capture_output=True, text=True)
    if r.returncode != 0 and san:
        pytest.skip("sanitizer runtime not available: %s" % r.stderr[-300:])
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="halt_on_error=1")
    out = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip() == "ok"


@pytest.mark.parametrize("san", [["-fsanitize=address,undefined", "-fno-omit-frame-pointer"]])
def test_oracle_under_sanitizers(tmp_path, san):
    """The oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
    on small frames (tests/native/oracle_san.c): EM in both modes with and
    without convergence, supervoxel kNN with a NaN, marshalling, local
    log-likelihood."""
    root = os.path.dirname(HERE)
    exe = str(tmp_path / "oracle_san")
    r = subprocess.run(["gcc", "-std=c11", "-O1", "-g", "-fopenmp", "-ffp-contract=off"] + san +
                       ["-o", exe, os.path.join(root, "oracle", "oracle.c"),
                        os.path.join(HERE, "native", "oracle_san.c"), "-lm"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("sanitizer runtime not available: %s" % r.stderr[-300:])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="halt_on_error=1",
               OMP_NUM_THREADS="2")
    out = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip() == "ok"
