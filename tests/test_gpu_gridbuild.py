"""The grid build on the device (tn_grid_build) against the reference's
build_grid_cpu (src/v2m_grid.cpp:47-722): every array bit for bit on the fixture
cases -- label volumes with the default, anisotropic and optional stages, a user
sizing field, probability inputs smoothed and raw, a gray-scale input, bricks
that overflow -- and, at the large workload's size (256^3), the properties the
fixtures pin, then seeding and a relaxation on the device-built grid."""
import os

import numpy as np
import pytest

from conftest import ROOT, bits

pytestmark = pytest.mark.gpu

GOLD = os.path.join(ROOT, "tests", "golden", "gridbuild.npz")
GB = np.load(GOLD)
CASES = [str(c) for c in GB["cases"]]


@pytest.mark.parametrize("name", CASES)
def test_grid_build_bit_exact(name):
    from gridbuild_cases import inputs_of

    from trussnet_amd import gridbuild

    print(f"[grid build] case {name}", flush=True)
    info = {}
    G = gridbuild.build(**inputs_of(GB, name), device=0, info=info)
    p = name + "_"
    for k in ("bl_cnt", "bl_lab", "bl_slot", "phi", "h", "grade"):
        got, want = getattr(G, k), GB[p + k]
        assert got.shape == want.shape, (k, got.shape, want.shape)
        assert np.array_equal(bits(got) if got.dtype == np.float32 else got,
                              bits(want) if want.dtype == np.float32 else want), (name, k)
    si, sf = GB[p + "scalars_i"], GB[p + "scalars_f"]
    assert [info["R"], info["limit_sweeps"], info["overflow_bricks"], info["nslot"]] == [int(x) for x in si]
    assert np.array_equal(bits(np.array([info["hbase"], info["hmin"], info["hmax"], info["hcurv"]], np.float32)),
                          bits(sf))
    if (p + "gTW") in GB:
        assert np.array_equal(bits(G.gTW), bits(GB[p + "gTW"]))


def test_grid_build_large_properties():
    """256^3, six nested perturbed ellipsoid labels: brick sets, slot scan, the
    limiter's fixed point and the grades hold at full size; seeding and 20
    iterations run on the device-built grid."""
    from trussnet_amd import gridbuild, relax, volumes
    from trussnet_amd._abi import GridParams, RelaxParams

    print("[grid build] 256^3 nested ellipsoids", flush=True)
    lab = volumes.nested_ellipsoids(256, nlab=6, seed=0, axes=(20.0, 120.0), grf_sigma=4.0, grf_amp=3.0)
    info = {}
    G = gridbuild.build(lab, prm=GridParams(hbase=4.5, hmin=1.5, hmax=6.0), device=0, info=info)
    n = np.minimum(G.bl_cnt, 16)
    mixed = n > 1
    starts = np.cumsum(np.where(mixed, n, 0)) - np.where(mixed, n, 0)
    assert np.array_equal(G.bl_slot, np.where(mixed, starts, -1))
    assert G.nslot == info["nslot"] > 1000
    assert np.all((G.phi >= -1e-6) & (G.phi <= 1 + 1e-6))
    h = G.h.reshape(G.nz, G.ny, G.nx)
    assert h.min() >= np.float32(G.hmin) and h.max() <= np.float32(G.hmax)
    g = np.float32(0.3)
    for ax in range(3):
        d = np.diff(h, axis=ax)
        assert np.all(np.abs(d) <= g * np.float32(1.0) * (1 + 1e-6) + 1e-6)
    t = np.log(G.h.astype(np.float64) / G.hmin) / np.log(G.hmax / G.hmin)
    want = np.floor(np.clip(t, 0, 1) * 255 + 0.5)
    assert np.abs(G.grade.astype(int) - want).max() <= 1
    assert info["limit_sweeps"] >= 2
    print(f"[grid build] {info['nslot']} slots, {info['limit_sweeps']} sweeps; seeding", flush=True)
    seeds = relax.seed(G, RelaxParams(), device=0)
    assert seeds.n > 100000
    out, st = relax.relax(G, seeds, RelaxParams(max_iters=20, dptol=0.0), device=0)
    assert st["iters"] == 20 and np.all(np.isfinite(out.P))


@pytest.mark.parametrize("name", ["sectors_opts", "gray40"])
def test_device_pipeline_equals_host_round_trip(name):
    """Seeding and a relaxation plan straight from the device-built grid
    (tn_seed_from_grid, tn_plan_create_from_grid) give the bits of the same steps
    on the downloaded grid."""
    from gridbuild_cases import inputs_of

    from trussnet_amd import gridbuild, relax
    from trussnet_amd._abi import RelaxParams

    print(f"[device pipeline] case {name}", flush=True)
    kw = inputs_of(GB, name)
    G = gridbuild.build(**kw, device=0)
    res = gridbuild.build_resident(**kw, device=0)
    try:
        s_dev = res.seed(RelaxParams())
        s_host = relax.seed(G, RelaxParams(), device=0)
        assert s_dev.equal(s_host) and np.array_equal(s_dev.part3, s_host.part3)
        prm = RelaxParams(max_iters=40, dptol=0.0)
        with res.plan(s_dev, prm) as plan:
            st = plan.run()
            got = plan.download()
        want, st_h = relax.relax(G, s_host, prm, device=0)
        assert got.equal(want)
        assert st["iters"] == st_h["iters"] == 40 and st["rebuilds"] == st_h["rebuilds"]
    finally:
        res.close()
