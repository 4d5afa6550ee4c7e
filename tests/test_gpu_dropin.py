"""The reference drivers' flow through the C++ drop-in header
(tests/cpp/dropin_wpk.cpp, built by __graft_entry__.build())."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_cpp_dropin_driver():
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "dropin_wpk")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is not built (run __graft_entry__.build())")
    r = subprocess.run([exe, "16"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0 and "DROPIN OK" in r.stdout
