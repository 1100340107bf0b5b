"""Write the grid-build fixtures from the compiled reference.

Run in a container that holds the reference sources (it needs
oracle/_ref/libv2m_ref.so, built by ``make -C oracle ref``):

    python tests/golden/make_gridbuild.py

For each case it runs the reference's build_grid_cpu (src/v2m_grid.cpp:47-722)
with every GridParams field set as the case says, and stores, as data only, the
inputs (labels, probabilities, gray-scale intensity and thresholds, the user's
sizing field, voxel size, parameters) and the Grid: brick counts, labels and
slots, phi, h, grades, the gray-scale thresholds and widths, and the scalars
(R, limit sweeps, overflow bricks, slots, hbase, hmin, hmax, hcurv).
File: tests/golden/gridbuild.npz (numpy, compressed, no pickles).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import ref  # noqa: E402
from trussnet_amd import volumes  # noqa: E402
from trussnet_amd._abi import GridParams  # noqa: E402

sys.path.insert(0, os.path.dirname(HERE))
from gridbuild_cases import PARAM_KEYS, TPM_ARGS  # noqa: E402



def many_labels(n: int = 40, block: int = 3, nlab: int = 40, seed: int = 4) -> np.ndarray:
    """Blocks of `block` voxels with random labels from a spread-out set (up to
    60000): bricks that see more than 16 labels within the radius."""
    rng = np.random.default_rng(seed)
    pool = np.concatenate([[0], rng.choice(np.arange(1, 60000), nlab - 1, replace=False)]).astype(np.uint16)
    nb = (n + block - 1) // block
    coarse = pool[rng.integers(0, nlab, size=(nb, nb, nb))]
    lab = np.repeat(np.repeat(np.repeat(coarse, block, 0), block, 1), block, 2)[:n, :n, :n]
    c = (n - 1) / 2.0
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    lab[(x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2 > (0.45 * n) ** 2] = 0
    return np.ascontiguousarray(lab)


def graded(lab, lo, hi, slope):
    from trussnet_amd.workload import graded_size
    return graded_size(lab, lo, hi, slope)


CASES = {
    # two concentric balls, the default parameters (thin-layer blend on)
    "spheres40": dict(vol=lambda: volumes.two_spheres(40, 16.0, 8.0), prm=GridParams(hbase=3.0)),
    # four labels with anisotropic voxels, label-preserving margin, thin-layer
    # sizing, per-label and interface sizes
    "sectors_opts": dict(vol=lambda: volumes.sectors(40), vs=(0.8, 0.9, 1.3),
                         prm=GridParams(hbase=2.5, preserve=0.05, thick=1.0, hlab=(0.0, 2.0, 0.0, 1.5, 0.0),
                                        iglob=2.2, ilab=(0.0, 0.0, 1.8), ipair=(1.0, 3.0, 1.2))),
    # perturbed ellipsoids with a user sizing field and no blend
    "ellipsoids_hvox": dict(vol=lambda: volumes.nested_ellipsoids(48, nlab=4, seed=3, axes=(6.0, 20.0), grf_sigma=2.0,
                                                                  grf_amp=1.0, centre_jitter=2.0),
                            hvox=(1.5, 4.0, 0.25), prm=GridParams(hbase=4.0, hmin=1.5, sigma_thin=0.0, K=2.5, g=0.25)),
    # a tissue-probability input, smoothed (sigma 1) and raw (sigma 0)
    "tpm_smooth": dict(vol=lambda: volumes.tpm_softmax(*TPM_ARGS)[1], prob=lambda: volumes.tpm_softmax(*TPM_ARGS)[0],
                       prm=GridParams(hbase=3.0, hmin=2.0)),
    "tpm_raw": dict(vol=lambda: volumes.tpm_softmax(*TPM_ARGS)[1], prob=lambda: volumes.tpm_softmax(*TPM_ARGS)[0],
                    prm=GridParams(hbase=3.0, sigma=0.0)),
    # a gray-scale input: memberships of the intensity
    "gray40": dict(gray=lambda: volumes.gray_ball(40), thr=volumes.GRAY_THRESHOLDS, prm=GridParams(hbase=2.5)),
    # more than 16 labels near some bricks, labels up to 60000, a sharper blend
    "many_labels": dict(vol=lambda: many_labels(30), prm=GridParams(hbase=3.0, sigma_thin=0.5, sigma_curv=1.5)),
}


def case_inputs(spec):
    """(labels, vs, prm, prob, gray, thresholds, hvox) of a case."""
    gray = thr = None
    if "gray" in spec:
        gray = spec["gray"]()
        thr = np.sort(np.asarray(spec["thr"], np.float32))
        lab = volumes.gray_labels(gray, thr)
    else:
        lab = spec["vol"]()
    prob = spec["prob"]() if "prob" in spec else None
    hvox = None
    if "hvox" in spec:
        lo, hi, slope = spec["hvox"]
        hvox = graded(lab, lo, hi, slope)
    return lab, spec.get("vs", (1.0, 1.0, 1.0)), spec["prm"], prob, gray, thr, hvox


def main(argv):
    names = argv or list(CASES)
    path = os.path.join(HERE, "gridbuild.npz")
    out = dict(np.load(path)) if argv and os.path.exists(path) else {}
    for name in names:
        lab, vs, prm, prob, gray, thr, hvox = case_inputs(CASES[name])
        r = ref.RefGrid.build_full(lab, vs, prm, prob=prob, gray=gray, thresholds=thr, hvox=hvox)
        G = r.grid()
        sc = r.scalars()
        p = name + "_"
        out[p + "labels"] = lab
        out[p + "vs"] = np.array(vs, np.float32)
        out[p + "params"] = np.array([getattr(prm, k) for k in PARAM_KEYS], np.float32)
        out[p + "hlab"] = np.array(prm.hlab, np.float32)
        out[p + "ilab"] = np.array(prm.ilab, np.float32)
        out[p + "ipair"] = np.array(prm.ipair, np.float32)
        # (the probabilities are not stored: TPM_ARGS regenerates them with
        # trussnet_amd.volumes.tpm_softmax)
        for k, a in (("gray", gray), ("thr", thr), ("hvox", hvox)):
            if a is not None:
                out[p + k] = np.asarray(a, np.float32)
        for k in ("bl_cnt", "bl_lab", "bl_slot", "phi", "h", "grade"):
            out[p + k] = getattr(G, k)
        if G.gm:
            out[p + "gTW"] = G.gTW
        out[p + "scalars_i"] = np.array([sc[k] for k in ("R", "limit_sweeps", "overflow_bricks", "nslot")], np.int64)
        out[p + "scalars_f"] = np.array([sc[k] for k in ("hbase", "hmin", "hmax", "hcurv")], np.float32)
        print(f"{name}: {lab.shape}, {sc['nslot']} slots, {sc['limit_sweeps']} sweeps, "
              f"{sc['overflow_bricks']} overflow bricks, h in [{sc['hmin']:.3g}, {sc['hmax']:.3g}]")
    out["cases"] = np.array(sorted({k[:-len("_labels")] for k in out if k.endswith("_labels")}))
    np.savez_compressed(path, **out)
    print(f"{path}: {os.path.getsize(path) / 1e3:.0f} kB")


if __name__ == "__main__":
    main(sys.argv[1:])
