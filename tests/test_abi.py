"""The C-ABI library loads, exports every symbol include/hipfoam.h declares,
and behaves on a host without a usable device (no compute calls)."""

import ctypes as C
import subprocess

import numpy as np

from metal4foam_amd import _lib
from metal4foam_amd import hipfoam as H


def test_header_and_binding_agree():
    assert _lib.header_symbols() == sorted(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    for name in _lib.header_symbols():
        assert name in exported, name
        assert getattr(lib, name) is not None


def test_library_is_gfx950_code_object():
    # the offload bundle embedded in the library names its target
    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_null_context_is_rejected():
    lib = _lib.load()
    assert lib.hipfoam_good(None) == 0
    assert lib.hipfoam_set_topology(None, 1, 0, None, None, None, None, None) == _lib.ERR_ARG
    st = _lib.Stats()
    assert lib.hipfoam_pcg(None, None, 1.0, 1, 1, 1, 2.0, 30.0, C.byref(st)) == _lib.ERR_ARG


def test_default_controls_are_reference_defaults():
    c = H.default_controls()
    assert (c.tolerance, c.relTol, c.maxIter, c.minIter) == (1e-6, 0.0, 1000, 0)
    assert (c.deflate, c.chebyshevOrder, c.chebyshevRatio) == (2, 8, 30.0)
    assert (c.maxRestarts, c.iterationsPerSync) == (8, 1)


def test_bad_device_gives_context_that_is_not_good():
    ctx = H.Context(device=999, require=False)
    assert not ctx.good
    assert "999" in ctx.description
    m_n = 4
    i32 = np.zeros(m_n + 1, np.int32)
    p = i32.ctypes.data_as(C.POINTER(C.c_int32))
    rc = ctx.lib.hipfoam_set_topology(ctx.h, m_n, 0, None, None, p, None, p)
    assert rc == _lib.ERR_DEVICE
    assert ctx.last_error
    ctx.close()


def test_schedule_validation_without_device():
    ctx = H.Context(device=999, require=False)
    s = ctx.get_schedule()
    assert (s.spmvSchedule, s.blockSize, s.xcdRemap, s.graphChunk) == (2, 0, 2, -1)
    bad = _lib.Schedule(3, 256, 1, 8, 0)
    assert ctx.lib.hipfoam_set_schedule(ctx.h, C.byref(bad)) == _lib.ERR_ARG
    bad = _lib.Schedule(0, 100, 1, 8, 0)
    assert ctx.lib.hipfoam_set_schedule(ctx.h, C.byref(bad)) == _lib.ERR_ARG
    bad = _lib.Schedule(0, 256, 1, 7, 0)
    assert ctx.lib.hipfoam_set_schedule(ctx.h, C.byref(bad)) == _lib.ERR_ARG
    bad = _lib.Schedule(0, 256, 3, 8, 0)
    assert ctx.lib.hipfoam_set_schedule(ctx.h, C.byref(bad)) == _lib.ERR_ARG
    for p in (-2, 2):
        bad = _lib.Schedule(0, 256, 1, 8, 0, 0, p)
        assert ctx.lib.hipfoam_set_schedule(ctx.h, C.byref(bad)) == _lib.ERR_ARG
    assert s.persistent == -1
    for lp in (-2, 1):
        bad = _lib.Schedule(0, 256, 1, 8, 0, 0, -1, lp)
        assert ctx.lib.hipfoam_set_schedule(ctx.h, C.byref(bad)) == _lib.ERR_ARG
    assert s.localLoop == -1
    assert ctx.local_launches() == 0
    assert ctx.lib.hipfoam_local_stats(ctx.h, None) == _lib.ERR_ARG
    assert ctx.lib.hipfoam_local_stats(None, C.byref(C.c_int64())) == _lib.ERR_ARG
    assert ctx.lib.hipfoam_simulate_abandon(None, 1) == _lib.ERR_ARG
    assert ctx.lib.hipfoam_simulate_abandon(ctx.h, -1) == _lib.ERR_ARG
    assert ctx.lib.hipfoam_simulate_abandon(ctx.h, 0) == _lib.OK
    ok = _lib.Schedule(1, 512, 0, 16, 1024)
    assert ctx.lib.hipfoam_set_schedule(ctx.h, C.byref(ok)) == _lib.OK
    assert ctx.get_schedule().as_dict() == ok.as_dict()
    # no mesh: no row-slot layout
    assert ctx.slot_layout() == (0, 0)
    rows = C.c_int64()
    assert ctx.lib.hipfoam_slot_layout(ctx.h, C.byref(rows), None) == _lib.ERR_ARG
    assert ctx.lib.hipfoam_slot_layout(None, C.byref(rows), C.byref(C.c_int32())) == _lib.ERR_ARG
    ctx.close()


def test_solver_selection_mirrors_run_time_tables():
    from metal4foam_amd import meshgen as mg
    sym = mg.poisson(3, 3, 1, seed=0)
    asym = mg.convection_diffusion(3, 3, 1, seed=0)
    assert isinstance(H.New("p", sym, {"solver": "hipPCG"}), H.hipPCG)
    assert isinstance(H.New("U", asym, {"solver": "hipPBiCGStab"}), H.hipPBiCGStab)
    assert isinstance(H.New("p", sym, {"solver": "hipPBiCGStab"}), H.hipPBiCGStab)
    for bad in ({"solver": "hipPCG"}, {"solver": "nope"}):
        try:
            H.New("U", asym, bad)
        except H.HipFoamError:
            pass
        else:
            raise AssertionError(bad)


def test_coupled_interfaces_are_rejected():
    from metal4foam_amd import meshgen as mg
    m = mg.poisson(3, 3, 1, seed=0)
    s = H.hipPCG("p", m, interfaces=[object()], solverControls={"solver": "hipPCG"})
    try:
        s.solve(np.zeros(m.n), m.source)
    except H.HipFoamError as e:
        assert "coupled" in str(e)
    else:
        raise AssertionError("coupled interface accepted")
