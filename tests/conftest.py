import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = ["spheres64", "jitter64", "sectors48", "ellipsoids56", "aniso40", "tpm32", "gray40", "shapes40"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: longer CPU test")


class Fixture:
    """One golden case written by tests/golden/make_golden.py from the compiled reference."""

    def __init__(self, name):
        from trussnet_amd._abi import Grid, Nodes

        self.name = name
        self.d = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
        d = self.d
        self.grid = Grid.from_arrays(d, "grid_")
        self.seeds = Nodes(d["seed_P"], d["seed_lab"], d["seed_typ"], d["seed_part"])
        self.stat_keys = [str(k) for k in d["stat_keys"]]

    def runs(self):
        from trussnet_amd._abi import Nodes, RelaxParams

        out = []
        for r, run in enumerate(self.d["runs"]):
            iters, dptol, fire = run[:3]
            vox = len(run) > 3 and bool(run[3])
            prm = RelaxParams(max_iters=int(iters), dptol=float(dptol), fire=bool(fire), voxel_trap=vox)
            res = Nodes(self.d[f"run{r}_P"], self.seeds.lab, self.d[f"run{r}_typ"], self.d[f"run{r}_part"])
            stats = dict(zip(self.stat_keys, self.d[f"run{r}_stats"]))
            out.append((prm, res, stats))
        return out


@pytest.fixture(scope="session", params=CASES)
def golden(request):
    return Fixture(request.param)


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import orc

    try:
        orc.lib()
    except RuntimeError:
        import subprocess

        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
    return orc


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a
