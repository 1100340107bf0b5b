"""bench.py's host-side helpers: the trace names of the kernels matched to the timed
classes, and the algorithmic bytes per launch (no device needed)."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_kernel_names_match_the_classes():
    b = _bench()
    names = {
        "void tnk::k_move<0, true>(tnd::Field, tnk::MoveArgs, tnk::Ctl*)": "k_move<0>",
        "void tnk::k_move<2, false>(tnd::Field)": "k_move<2>",
        "void tnk::k_move_easy<true>(tnd::Field, tnk::MoveArgs)": "k_move_easy",
        "tnk::k_truss(tnd::Hash, tnd::TrussIn, int, int*, int*)": "k_truss",
        "tnk::k_force(HIP_vector_type<float, 4u> const*, unsigned char const*)": "k_force",
        "__amd_rocclr_copyBuffer": "__amd_rocclr_copyBuffer",
    }
    for full, key in names.items():
        assert b.kernel_key(full) == key
    # every kernel a class names is one kernel_key can produce
    for ks in b.CLASS_KERNELS.values():
        for k in ks:
            assert b.kernel_key(f"void tnk::{k}(int)") == k


def test_launch_bytes_scale_with_items():
    b = _bench()
    per = b.item_bytes(21.0, 3)
    t = {"neighbors": (2, 1.0, 2 * 1000.0), "force": (4, 1.0, 4 * 1000.0)}
    assert b.launch_bytes("neighbors", t, per) == per["neighbors"] * 1000.0
    assert b.launch_bytes("force", t, per) == per["force"] * 1000.0
    assert b.launch_bytes("force", {"force": (0, 0.0, 0.0)}, per) is None


def test_classes_of_one_kernel_merge():
    b = _bench()
    per = b.item_bytes(21.0, 3)
    t = {"move_rest": (3, 1.5, 3 * 4000.0), "move_def": (3, 0.3, 3 * 30.0), "force": (3, 0.5, 3 * 1e6),
         "control": (3, 0.1, 3.0)}
    k = b.by_kernel(t, per)
    assert set(k) == {"k_move<0>", "k_force"}
    m = k["k_move<0>"]
    assert m["classes"] == ["move_rest", "move_def"] and m["launches"] == 6
    assert abs(m["ms"] - 1.8) < 1e-12 and m["items"] == 3 * 4030.0
    assert abs(m["bytes"] - per["move_rest"] * 3 * 4000.0 - per["move_def"] * 3 * 30.0) < 1e-6
