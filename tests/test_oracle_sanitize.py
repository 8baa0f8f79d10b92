"""The oracle's C sources under AddressSanitizer and UndefinedBehaviorSanitizer (host only).

`make -C oracle sanitize` builds the same sources with `-fsanitize=address,undefined`
(every UB report fatal); tests/oracle_sanitize_run.py then calls every C entry point
(Nedelec-2 element matrices in both frames, CSC pattern/fill/scatter, point location,
interpolation, Robin blocks and excitations, heat matrices) on a hex and a Delaunay mesh
in a child process with the ASan runtime preloaded. The run must end cleanly, print no
sanitizer report, and give the normal build's results: integer outputs equal, values
within rounding (the sanitizer build schedules the fast-math arithmetic differently).
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


def _runtime(name):
    p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True,
                       text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def _run(out, env_extra):
    env = dict(os.environ, OMP_NUM_THREADS="2", **env_extra)
    if "EMERGE_ORACLE_LIB" not in env_extra:
        env.pop("EMERGE_ORACLE_LIB", None)
    return subprocess.run([sys.executable, os.path.join(HERE, "oracle_sanitize_run.py"), out],
                          env=env, capture_output=True, text=True, timeout=600)


@pytest.mark.skipif(shutil.which("gcc") is None or _runtime("libasan.so") is None,
                    reason="gcc with the AddressSanitizer runtime is needed")
def test_oracle_under_asan_ubsan(tmp_path):
    san = tmp_path / "san"
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "sanitize",
                    f"SAN_OUT={san}"], check=True)
    lib = san / "libned2_oracle_san.so"
    syms = subprocess.run(["nm", "-D", str(lib)], capture_output=True, text=True).stdout
    assert "__asan_report" in syms and "__ubsan_handle" in syms  # instrumented
    plain = _run(str(tmp_path / "plain.npz"), {})
    assert plain.returncode == 0, plain.stderr[-4000:]
    checked = _run(str(tmp_path / "san.npz"), {
        "EMERGE_ORACLE_LIB": str(lib),
        "LD_PRELOAD": _runtime("libasan.so"),
        "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
    })
    report = checked.stderr
    assert checked.returncode == 0, report[-4000:]
    assert f"oracle library: {lib}" in checked.stdout  # the instrumented build ran
    assert "AddressSanitizer" not in report and "runtime error" not in report, report[-4000:]
    a, b = np.load(tmp_path / "plain.npz"), np.load(tmp_path / "san.npz")
    assert sorted(a.files) == sorted(b.files) and len(a.files) > 30
    for k in a.files:
        x, y = a[k], b[k]
        assert x.shape == y.shape and x.dtype == y.dtype, k
        if np.issubdtype(x.dtype, np.integer):
            assert np.array_equal(x, y), k
        else:
            assert np.array_equal(np.isnan(x), np.isnan(y)), k
            fin = ~np.isnan(x)
            scale = np.abs(x[fin]).max() if fin.any() else 1.0
            assert np.abs(x[fin] - y[fin]).max() <= 1e-11 * scale, k
