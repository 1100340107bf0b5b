"""The inputs of the grid-build fixtures (tests/golden/gridbuild.npz) as the device
build takes them: labels, voxel size, GridParams, and the optional probability,
gray-scale and sizing inputs."""
from __future__ import annotations

import numpy as np

from trussnet_amd import volumes
from trussnet_amd._abi import GridParams

TPM_ARGS = ((32, 34, 30), 6, 3.0, 1)   # tpm_softmax(shape, channels, sigma, seed) of the tpm_ cases
PARAM_KEYS = ["sigma", "sigma_thin", "thin_lo", "thin_hi", "thick", "thin_floor", "preserve", "sigma_curv", "hbase",
              "hmin", "hmax", "K", "g", "iglob"]


def params_of(gb, name: str) -> GridParams:
    p = name + "_"
    kw = {k: float(v) for k, v in zip(PARAM_KEYS, gb[p + "params"])}
    return GridParams(**kw, hlab=tuple(gb[p + "hlab"]), ilab=tuple(gb[p + "ilab"]), ipair=tuple(gb[p + "ipair"]))


def inputs_of(gb, name: str) -> dict:
    """Keyword arguments of trussnet_amd.gridbuild.build for a fixture case."""
    p = name + "_"
    kw = dict(labels=gb[p + "labels"], vs=tuple(float(v) for v in gb[p + "vs"]), prm=params_of(gb, name))
    if name.startswith("tpm_"):
        kw["prob"] = volumes.tpm_softmax(*TPM_ARGS)[0]
    if (p + "gray") in gb:
        kw["gray"], kw["thresholds"] = gb[p + "gray"], gb[p + "thr"]
    if (p + "hvox") in gb:
        kw["hvox"] = gb[p + "hvox"]
    return kw
