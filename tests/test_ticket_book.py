"""The serving loop's ticket book (hunyuansc25_amd/csrc/hgp_tickets.hpp) on
the host alone: built with g++ (with AddressSanitizer and
UndefinedBehaviorSanitizer) from tests/native/ticket_book_test.cpp and run.
It checks that a ticket is waited for once (a racing second waiter is
refused without touching the job) and that the teardown drain returns only
after every wait in flight."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_ticket_book_threads(tmp_path):
    exe = str(tmp_path / "ticket_book_test")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-pthread", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-I", os.path.join(ROOT, "hunyuansc25_amd", "csrc"),
                    os.path.join(HERE, "native", "ticket_book_test.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "ok"
