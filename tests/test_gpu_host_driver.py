"""The C++ drop-in classes (eol-cloth_amd/host/eolc_dropin.hpp) driven by the host
step loop host/cloth_step: 10 steps of the 20x20 cloth against the oracle's
Cloth::step, positions within 1e-10 (relative to max |x|)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import GRAV, H, rel_err
from eolc import scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "eol-cloth_amd", "host", "cloth_step")


@pytest.mark.gpu
def test_cpp_dropin_step_loop(tmp_path):
    out = tmp_path / "x.bin"
    r = subprocess.run([EXE, "20", "10", str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    xg = np.fromfile(out, dtype=np.float64)
    sc = scene.Scene(20)
    x, v = sc.x0.copy(), sc.v0.copy()
    for _ in range(10):
        rc, st, _ = O.step(x, v, sc.X, sc.tri, sc.hinge, O.material(), GRAV, H, sc.fixed, tol=1e-12)
        assert rc == 0
    assert rel_err(xg, x) < 1e-10


def test_cpp_dropin_driver_is_built():
    assert os.path.exists(EXE), "run __graft_entry__.build()"
