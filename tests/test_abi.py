"""The C ABI library: it loads, exports every declared entry point, its structs
match the header, and without a device every call fails with a status (the
caller's cue to fall back, as src/v2m_pipeline.cpp:497-526 does)."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, Fixture

HEADER = os.path.join(ROOT, "include", "trussnet_amd.h")


@pytest.fixture(scope="module")
def tnlib():
    from trussnet_amd import relax

    try:
        relax.lib()
    except RuntimeError:
        from trussnet_amd import build

        build.build()
    return relax


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tn_[a-z_]+)\s*\(", text)))


def test_exports_every_declared_symbol(tnlib):
    names = declared_functions()
    assert len(names) >= 13
    out = subprocess.run(["nm", "-D", "--defined-only", tnlib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(tn_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    assert sorted(tnlib.SYMBOLS) == names


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors and the C header agree on every struct size and offset."""
    import ctypes as C

    from trussnet_amd import _abi
    from trussnet_amd.adjoint import AdjProblemC

    structs = {"tn_grid_view": _abi.GridView, "tn_relax_params": _abi.RelaxParamsC, "tn_nodes_view": _abi.NodesView,
               "tn_relax_stats": _abi.RelaxStatsC, "tn_hash_info": _abi.HashInfoC,
               "tn_kernel_time": _abi.KernelTimeC, "tn_adj_problem": AdjProblemC,
               "tn_grid_params": _abi.GridParamsC, "tn_volume_view": _abi.VolumeViewC,
               "tn_grid_info": _abi.GridInfoC}
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, cls in structs.items():
        src.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            src.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    src.append("return 0; }")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(c)], check=True)
    lines = dict(ln.split() for ln in subprocess.run([str(exe)], capture_output=True, text=True,
                                                     check=True).stdout.splitlines())
    for cname, cls in structs.items():
        assert int(lines[cname]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(lines[f"{cname}.{fname}"]) == getattr(cls, fname).offset, (cname, fname)


def test_default_params_match_reference(tnlib):
    from dataclasses import fields

    from trussnet_amd._abi import RelaxParams

    got, want = tnlib.default_params(), RelaxParams()
    for f in fields(RelaxParams):
        assert np.float32(getattr(got, f.name)) == np.float32(getattr(want, f.name)), f.name


def test_default_grid_params_match_reference(tnlib):
    """tn_grid_params_default is GridParams' defaults (src/v2m_grid.h:21-43)."""
    import ctypes as C

    from trussnet_amd import gridbuild
    from trussnet_amd._abi import GridParams, GridParamsC

    L = gridbuild._bind()
    got = GridParamsC()
    L.tn_grid_params_default(C.cast(C.byref(got), C.c_void_p))
    want, _ = GridParams().to_c()
    for name, _t in GridParamsC._fields_:
        if name in ("hlab", "ilab", "ipair"):
            assert not getattr(got, name), name
        else:
            assert getattr(got, name) == getattr(want, name), name


def test_grid_build_rejects_bad_volumes(tnlib):
    """Malformed volumes fail on the host before any device work."""
    from trussnet_amd import gridbuild, volumes
    from trussnet_amd._abi import GridParams

    lab = volumes.two_spheres(16, 6.0, 3.0)
    gray = np.zeros(lab.shape, np.float32)
    with pytest.raises(tnlib.DeviceError):   # descending thresholds
        gridbuild.build(lab, gray=gray, thresholds=None, prm=GridParams(sigma_curv=0.0))
    v, keep = gridbuild.volume_view(lab, gray=gray, thresholds=[0.2, 0.5])
    v.thresholds[0], v.thresholds[1] = 0.5, 0.2
    import ctypes as C

    pc, pk = GridParams().to_c()
    h = C.c_void_p()
    rc = gridbuild._bind().tn_grid_build(C.byref(v), C.cast(C.byref(pc), C.c_void_p), -1, C.byref(h))
    from trussnet_amd._abi import TN_ERR_ARG

    assert rc == TN_ERR_ARG


def test_shape_program_verified(tnlib):
    """A shape grid's program is walked on the host before any device work: every
    read the interpreter can make must lie inside it (a corrupt one fails with
    TN_ERR_ARG; a sound one gets as far as the device)."""
    from trussnet_amd._abi import TN_ERR_ARG, Grid

    fx = Fixture("shapes40")
    g = fx.grid
    assert g.gm < 0 and g.gTW.size > 100
    bad_cases = []
    p = g.gTW.copy()
    n = int(p[0])
    p[5] = p.size + 10          # label 0's code beyond the program
    bad_cases.append(p)
    p = g.gTW.copy()
    p[int(p[5])] = 77           # an unknown op
    bad_cases.append(p)
    p = g.gTW.copy()
    p[8 + n] = p.size - 2       # a brick table cut off
    bad_cases.append(p)
    for prog in bad_cases:
        bad = Grid(g.nx, g.ny, g.nz, g.vs, g.L, g.bl_cnt, g.bl_lab, g.bl_slot, g.phi, g.h, g.hmin, g.hmax, None,
                   None, prog, g.gm)
        with pytest.raises(tnlib.DeviceError) as e:
            tnlib.relax(bad, fx.seeds)
        assert f"[{TN_ERR_ARG}]" in str(e.value), str(e.value)
    if not os.path.exists("/dev/kfd"):
        with pytest.raises(tnlib.NoDevice):
            tnlib.relax(g, fx.seeds)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible here")
def test_no_device_raises(tnlib):
    """No HIP device: tn_relax fails with TN_ERR_NODEVICE and the nodes stay as they were."""
    fx = Fixture("spheres64")
    assert tnlib.device_count() == 0
    with pytest.raises(tnlib.NoDevice):
        tnlib.relax(fx.grid, fx.seeds)
    with pytest.raises(tnlib.NoDevice):
        tnlib.Plan(fx.grid, fx.seeds)


def test_bad_arguments_rejected(tnlib):
    """Malformed inputs fail on the host before any device work (any status but OK)."""
    from trussnet_amd._abi import Grid, RelaxParams

    fx = Fixture("spheres64")
    g = fx.grid
    bad = Grid(g.nx, g.ny, g.nz, g.vs, g.L, g.bl_cnt, g.bl_lab, g.bl_slot.copy(), g.phi, g.h, g.hmin, g.hmax)
    bad.bl_slot[np.argmax(bad.bl_slot)] = g.nslot + 5  # a slot beyond phi
    with pytest.raises(tnlib.DeviceError):
        tnlib.relax(bad, fx.seeds)
    with pytest.raises(tnlib.DeviceError):
        tnlib.relax(g, fx.seeds, RelaxParams(max_iters=-1))


def test_cpp_wrapper_builds_and_links(tnlib, tmp_path):
    """include/trussnet_amd.hpp compiles as C++17 and links against the library."""
    src = tmp_path / "use.cpp"
    src.write_text(
        '#include "trussnet_amd.hpp"\n#include <cstdio>\n'
        "int main() {\n"
        "  tn_relax_params p = tnamd::default_params();\n"
        "  tn_grid_params gp = tnamd::default_grid_params();\n"
        "  tn_volume_view v{};\n"
        "  try { tnamd::DeviceGrid g(v, gp); } catch (const tnamd::Error& e) { std::printf(\"%d \", e.code()); }\n"
        "  try { tnamd::check(TN_ERR_ARG, \"probe\"); } catch (const std::runtime_error& e) {\n"
        "    std::printf(\"%.3f %d %d %.1f\\n\", p.t, p.max_iters, static_cast<const tnamd::Error&>(e).code(),\n"
        "                gp.sigma_curv);\n"
        "  }\n  return 0;\n}\n")
    exe = tmp_path / "use"
    libdir = os.path.dirname(tnlib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe), f"-L{libdir}", "-ltrussnet_amd", f"-Wl,-rpath,{libdir}"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert out == [str(1), "1.300", "500", str(1), "2.0"]   # an empty volume is refused on the host
