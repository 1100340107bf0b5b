"""The v2mesh-side binding on the device: oracle/_ref/check_v2m_hip (built from
trussnet_amd/integration/v2m_hip.cpp and tests/integration/check_v2m_hip.cpp
against the reference's headers and CPU path, where the reference sources are)
runs build_grid_hip, seed_hip, relax_hip and relax_hip_batch through the
reference's own Grid / Nodes / RelaxStats types and compares each with
build_grid_cpu (prob_fields included), seed_cpu and relax_cpu, bit for bit."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "check_v2m_hip")


def test_binding_on_the_device():
    if not os.path.exists(BIN):
        pytest.skip("oracle/_ref/check_v2m_hip is built only where the reference sources are")
    print("[integration] check_v2m_hip device", flush=True)
    r = subprocess.run([BIN, "device"], capture_output=True, text=True, timeout=300)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().splitlines()[-1] == "ok device", r.stdout
