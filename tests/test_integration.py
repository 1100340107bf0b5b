"""The v2mesh-side binding (trussnet_amd/integration/v2m_hip.cpp) against the
reference's own headers and CPU path.

`make -C oracle integration` compiles the binding and its check program
(tests/integration/check_v2m_hip.cpp) with the reference's src/ headers and links
them with the reference's compiled CPU path (oracle/_ref/libv2m_ref.so) and
libtrussnet_amd. Without a device every *_hip call must throw
std::runtime_error and the pipeline's fallback (restore the seeds, run relax_cpu,
src/v2m_pipeline.cpp:497-526) must give relax_cpu's nodes. Needs the reference
sources (this container); the GPU test runs the same program on the device.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"
BIN = os.path.join(ROOT, "oracle", "_ref", "check_v2m_hip")
LIB = os.path.join(ROOT, "trussnet_amd", "lib", "libtrussnet_amd.so")


@pytest.mark.skipif(not os.path.isdir(REF), reason="needs the reference sources to compile against")
def test_binding_compiles_and_falls_back_without_a_device():
    if not os.path.exists(LIB):
        pytest.skip("libtrussnet_amd.so not built")
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "integration"], check=True, capture_output=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300, env={**os.environ, "HIP_VISIBLE_DEVICES": ""})
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().splitlines()[-1] == "ok fallback", r.stdout
