"""The C++ mirror of Simulation/ObjectAPI drives a pile on the GPU (host/example_pile.cpp)."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.gpu
def test_host_mirror_example_runs():
    exe = ROOT / "physicssimulation-agony_amd" / "host" / "example_pile"
    if not exe.exists():
        pytest.fail("host/example_pile is not built (run __graft_entry__.build())")
    print("running", exe, flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "manifolds" in r.stdout
