"""The library's host-only code (generators, graph-file reader and writers,
the host CSR check) and the oracle, built with AddressSanitizer and
UndefinedBehaviorSanitizer by tests/native/Makefile and run over generated
graphs and the committed graph files (CPU only)."""
import glob
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="no host compiler")
def test_host_code_under_sanitizers(tmp_path):
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "native"), "sanitize", f"OUT={tmp_path}"], check=True)
    files = sorted(glob.glob(os.path.join(HERE, "golden", "io", "*.graph")) +
                   glob.glob(os.path.join(HERE, "golden", "io", "*.mtx")))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(tmp_path / "host_sanitize"), str(tmp_path)] + files, capture_output=True, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert r.stdout.strip() == "ok"
