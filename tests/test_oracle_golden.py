"""The CPU restatement (oracle/oracle.c) against the reference's own outputs.

Every value here was written by the compiled reference (tests/golden/make_golden.py):
the restatement must reproduce each bit for bit, stage by stage and over whole
relaxations (src/v2m_particles.cpp:728-976).
"""
import numpy as np
import pytest

from conftest import bits
from trussnet_amd._abi import K, RelaxParams


def test_hash_geometry(golden, oracle_lib):
    hi = oracle_lib.hash_info(golden.grid, RelaxParams())
    d = golden.d
    assert hi.nlev == int(d["hash_nlev"]) and hi.nkeys == int(d["hash_nkeys"])
    assert np.float32(hi.b0) == d["hash_b0"]
    assert [list(hi.dim[L]) for L in range(8)] == d["hash_dim"].tolist()
    assert list(hi.off) == d["hash_off"].tolist()


def test_rebuild_bit_exact(golden, oracle_lib):
    d = golden.d
    rb = oracle_lib.rebuild(golden.grid, golden.seeds, RelaxParams())
    for k in ("key", "sorted", "cstart", "nnb"):
        assert np.array_equal(rb[k], d["rb_" + k]), k
    assert np.array_equal(bits(rb["hn"]), bits(d["rb_hn"]))
    nbr = rb["nbr"].reshape(-1, K).copy()
    nbr[np.arange(K)[None, :] >= rb["nnb"][:, None]] = -1
    assert np.array_equal(nbr.ravel(), d["rb_nbr"])


def test_force_bit_exact(golden, oracle_lib):
    d = golden.d
    nbr = d["rb_nbr"].copy()
    nbr[nbr < 0] = 0
    F = oracle_lib.force(golden.seeds, RelaxParams(), d["rb_hn"], nbr, d["rb_nnb"])
    assert np.array_equal(bits(F), bits(d["force_F"]))


def test_move_from_rest_bit_exact(golden, oracle_lib):
    d = golden.d
    V0 = np.zeros(golden.seeds.n * 3, np.float32)
    out, hn, V, pw, mv = oracle_lib.move(golden.grid, golden.seeds, RelaxParams(), d["rb_hn"], d["force_F"], V0,
                                         float(d["move_fdt"]), 0.1)
    assert np.array_equal(bits(out.P), bits(d["move_P"]))
    assert np.array_equal(out.typ, d["move_typ"]) and np.array_equal(out.part, d["move_part"])
    for name, a in (("hn", hn), ("V", V), ("pw", pw), ("mv", mv)):
        assert np.array_equal(bits(a), bits(d["move_" + name])), name


def test_field_queries_bit_exact(golden, oracle_lib):
    d = golden.d
    a, b = (int(x) for x in d["fq_ab"])
    phi, psi, grad, lab, sec = oracle_lib.field_points(golden.grid, a, b, d["fq_X"])
    assert np.array_equal(bits(phi), bits(d["fq_phi"]))
    assert np.array_equal(bits(psi), bits(d["fq_psi"]))
    assert np.array_equal(bits(grad), bits(d["fq_grad"]))
    assert np.array_equal(lab, d["fq_label"]) and np.array_equal(sec, d["fq_second"])


def test_relax_runs_bit_exact(golden, oracle_lib):
    for prm, want, stats in golden.runs():
        got, st = oracle_lib.relax(golden.grid, golden.seeds, prm)
        assert got.equal(want), (golden.name, prm.max_iters, prm.dptol, prm.fire)
        for k, v in stats.items():
            assert np.float32(st[k]) == np.float32(v), (golden.name, prm.max_iters, k, st[k], v)


@pytest.mark.parametrize("nthreads", [1, 3])
def test_relax_thread_count_invariant(oracle_lib, nthreads):
    """The reference's results do not depend on the OpenMP thread count; nor do the oracle's."""
    from conftest import Fixture

    fx = Fixture("sectors48")
    prm, want, _ = fx.runs()[1]
    got, _ = oracle_lib.relax(fx.grid, fx.seeds, prm, nthreads=nthreads)
    assert got.equal(want)
