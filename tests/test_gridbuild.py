"""The grid-build fixtures (tests/golden/gridbuild.npz, the reference's
build_grid_cpu, src/v2m_grid.cpp:47-722): they are what the compiled reference
makes today, and they hold what the mathematics of each stage fixes -- the brick
label sets, the slot scan, the gradient limit's fixed point, the grades, the
gray-scale widths and memberships, the raw probability fields -- so that the
device build is checked against pinned data."""
import os

import numpy as np
import pytest

from conftest import ROOT

GOLD = os.path.join(ROOT, "tests", "golden", "gridbuild.npz")


@pytest.fixture(scope="module")
def gb():
    return np.load(GOLD)


def cases(gb):
    return [str(c) for c in gb["cases"]]


def grid_of(gb, name):
    from trussnet_amd._abi import Grid

    p = name + "_"
    lab = gb[p + "labels"]
    nz, ny, nx = lab.shape
    sf = gb[p + "scalars_f"]
    gm = gb[p + "gTW"].size // 2 if (p + "gTW") in gb else 0
    return Grid(nx, ny, nz, tuple(float(v) for v in gb[p + "vs"]), lab.ravel(), gb[p + "bl_cnt"], gb[p + "bl_lab"],
                gb[p + "bl_slot"], gb[p + "phi"], gb[p + "h"], float(sf[1]), float(sf[2]), gb[p + "grade"],
                gb[p + "gray"].ravel() if gm else None, gb[p + "gTW"] if gm else None, gm)


def test_fixture_is_the_reference(gb):
    """The compiled reference still makes every fixture bit for bit (where it is built)."""
    try:
        from oracle import ref

        ref.lib()
    except Exception as e:   # the reference is not built in this container
        pytest.skip(f"reference not built: {e}")
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_gridbuild", os.path.join(ROOT, "tests", "golden",
                                                                                 "make_gridbuild.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    for name in cases(gb):
        lab, vs, prm, prob, gray, thr, hvox = mg.case_inputs(mg.CASES[name])
        r = ref.RefGrid.build_full(lab, vs, prm, prob=prob, gray=gray, thresholds=thr, hvox=hvox)
        G = r.grid()
        for k in ("bl_cnt", "bl_lab", "bl_slot", "phi", "h", "grade"):
            assert np.array_equal(getattr(G, k).view(np.uint8), gb[name + "_" + k].view(np.uint8)), (name, k)


def _dilated_label_sets(lab, R, nbx, nby, nbz):
    """Brute force: the sorted labels within R of each brick (label 0 outside)."""
    nz, ny, nx = lab.shape
    pad = np.zeros((nz + 2 * R + 8, ny + 2 * R + 8, nx + 2 * R + 8), lab.dtype)
    pad[R:R + nz, R:R + ny, R:R + nx] = lab
    out = []
    for b in range(nbx * nby * nbz):
        bx, by, bz = b % nbx, (b // nbx) % nby, b // (nbx * nby)
        blk = pad[bz * 8:bz * 8 + 8 + 2 * R + 1, by * 8:by * 8 + 8 + 2 * R + 1, bx * 8:bx * 8 + 8 + 2 * R + 1]
        out.append(np.unique(blk))
    return out


def test_bricks_and_slots(gb):
    """bl_lab lists each brick's labels within R in ascending order (16 at most,
    count 17 on overflow, 0xFFFF padded); bl_slot is the exclusive scan of the
    mixed bricks' counts (src/v2m_grid.cpp:103-147)."""
    for name in cases(gb):
        g = grid_of(gb, name)
        R = int(gb[name + "_scalars_i"][0])
        sets = _dilated_label_sets(gb[name + "_labels"], R, g.nbx, g.nby, g.nbz)
        cnt, labs = g.bl_cnt, g.bl_lab.reshape(-1, 16)
        nover = 0
        for b, st in enumerate(sets):
            if st.size <= 16:
                assert cnt[b] == st.size, (name, b)
                assert np.array_equal(labs[b, :st.size], st), (name, b)
                assert np.all(labs[b, st.size:] == 0xFFFF)
            else:
                nover += 1
                assert cnt[b] == 17
                assert np.all(np.isin(labs[b], st)) and np.all(np.diff(labs[b].astype(int)) > 0)
        assert nover == int(gb[name + "_scalars_i"][2])
        n = np.minimum(cnt, 16)
        mixed = n > 1
        starts = np.cumsum(np.where(mixed, n, 0)) - np.where(mixed, n, 0)
        assert np.array_equal(g.bl_slot, np.where(mixed, starts, -1))
        assert g.nslot == int(np.where(mixed, n, 0).sum()) == int(gb[name + "_scalars_i"][3])
    assert int(gb["many_labels_scalars_i"][2]) > 0   # the overflow path is covered


def test_gradient_limit_fixed_point(gb):
    """h is a fixed point of the limiter: h <= h(n) + g |x - x_n| over the 26
    neighbours, and within [hmin, hmax] (src/opencl/v2m_grid_body.cl:477-496)."""
    from gridbuild_cases import params_of

    for name in cases(gb):
        g = grid_of(gb, name)
        prm = params_of(gb, name)
        h = g.h.reshape(g.nz, g.ny, g.nx).astype(np.float32)
        vs = np.array(g.vs, np.float32)
        assert h.min() >= np.float32(g.hmin) * (1 - 1e-6) and h.max() <= np.float32(g.hmax) * (1 + 1e-6)
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    if dx == dy == dz == 0:
                        continue
                    d = np.float32(prm.g) * np.sqrt(np.float32((dx * vs[0]) ** 2 + (dy * vs[1]) ** 2
                                                               + (dz * vs[2]) ** 2), dtype=np.float32)
                    a = h[max(0, dz):h.shape[0] + min(0, dz), max(0, dy):h.shape[1] + min(0, dy),
                          max(0, dx):h.shape[2] + min(0, dx)]
                    nb = h[max(0, -dz):h.shape[0] + min(0, -dz), max(0, -dy):h.shape[1] + min(0, -dy),
                           max(0, -dx):h.shape[2] + min(0, -dx)]
                    assert np.all(a <= nb + d), (name, dx, dy, dz)


def test_grades(gb):
    """grade = round(255 log(h / hmin) / log(hmax / hmin)), clamped (v2m_grade)."""
    for name in cases(gb):
        g = grid_of(gb, name)
        t = np.log(g.h.astype(np.float64) / g.hmin) / np.log(g.hmax / g.hmin)
        want = np.floor(np.clip(t, 0, 1) * 255 + 0.5)
        assert np.abs(g.grade.astype(int) - want).max() <= 1, name
        assert np.mean(g.grade == want) > 0.999, name


def test_gray_widths_and_memberships(gb):
    """W_k = 4 x the median intensity step across t_k between axis neighbours, and
    phi_l = clamp(0.5 + min((I - t_{l-1}) / W_{l-1}, (t_l - I) / W_l)) on every voxel
    of a slot (src/v2m_grid.cpp:374-433)."""
    g = grid_of(gb, "gray40")
    I = gb["gray40_gray"].astype(np.float32)
    t = gb["gray40_thr"].astype(np.float32)
    m = t.size
    W = g.gTW[m:]
    assert np.array_equal(g.gTW[:m], t)
    for k in range(m):
        steps = []
        for ax in range(3):
            a = np.moveaxis(I, 2 - ax, 0)
            x, y = a[:-1] - t[k], a[1:] - t[k]
            sel = (x < 0) != (y < 0)
            steps.append(np.abs(y - x)[sel])
        s = np.sort(np.concatenate(steps))
        assert W[k] == np.float32(4.0) * s[s.size // 2]
    slabs = g.bl_lab.reshape(-1, 16)
    for b in np.flatnonzero(g.bl_slot >= 0)[:200]:
        bx, by, bz = b % g.nbx, (b // g.nbx) % g.nby, b // (g.nbx * g.nby)
        blk = I[bz * 8:bz * 8 + 8, by * 8:by * 8 + 8, bx * 8:bx * 8 + 8]
        if blk.shape != (8, 8, 8):
            continue
        for s in range(min(int(g.bl_cnt[b]), 16)):
            l = int(slabs[b, s])
            mm = np.full(blk.shape, np.float32(1e30))
            if 1 <= l <= m:
                mm = np.minimum(mm, (blk - t[l - 1]) / W[l - 1])
            if l < m:
                mm = np.minimum(mm, (t[l] - blk) / W[l])
            want = np.clip(np.float32(0.5) + mm, 0, 1).astype(np.float32)
            got = g.phi[(g.bl_slot[b] + s) * 512:(g.bl_slot[b] + s + 1) * 512].reshape(8, 8, 8)
            assert np.array_equal(got, want), (b, l)


def test_raw_probability_fields(gb):
    """With sigma 0 a probability input's fields are the probabilities themselves."""
    from trussnet_amd import volumes

    from gridbuild_cases import TPM_ARGS

    g = grid_of(gb, "tpm_raw")
    prob = volumes.tpm_softmax(*TPM_ARGS)[0]
    slabs = g.bl_lab.reshape(-1, 16)
    for b in np.flatnonzero(g.bl_slot >= 0)[:100]:
        bx, by, bz = b % g.nbx, (b // g.nbx) % g.nby, b // (g.nbx * g.nby)
        for s in range(min(int(g.bl_cnt[b]), 16)):
            l = int(slabs[b, s])
            full = np.zeros((8, 8, 8), np.float32) + (1.0 if l == 0 else 0.0)
            src = prob[l, bz * 8:bz * 8 + 8, by * 8:by * 8 + 8, bx * 8:bx * 8 + 8]
            full[:src.shape[0], :src.shape[1], :src.shape[2]] = src
            got = g.phi[(g.bl_slot[b] + s) * 512:(g.bl_slot[b] + s + 1) * 512].reshape(8, 8, 8)
            assert np.array_equal(got, full), (b, l)
