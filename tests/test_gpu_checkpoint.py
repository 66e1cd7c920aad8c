"""Checkpoints and restart (camr_write_checkpoint, after CAMR::checkPoint,
Source/Utils/CAMR_io.cpp:45-84): a run restarted from a checkpoint
continues bit for bit like the uninterrupted run, for a two-level
hierarchy with flux registers and for a single fused-engine level; the
checkpoint header and state read back exactly."""
import numpy as np
import pytest

from camr_amd import _lib, workloads
from camr_amd.hydro import FluxRegister, Level, Problem, amr_step
from camr_amd.plotfile import read_checkpoint, restart, write_checkpoint
from tests.test_oracle_amr import FINE2, _boxes, _field

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _level(n, boxes):
    prob = Problem(domlo=(0, 0, 0), domhi=(n - 1,) * 3, dx=(1.0 / n,) * 3, bclo=(0,) * 3,
                   bchi=(0,) * 3, is_periodic=(1, 1, 1), boxes=boxes,
                   params=dict(ppm_type=1, cfl=0.3))
    return Level(prob, engine=_lib.ENGINE_STAGED)


def _hierarchy(init=True):
    C, F = _level(16, _boxes(16, 8)), _level(32, FINE2)
    if init:
        for L, n in ((C, 16), (F, 32)):
            for b, bb in enumerate(L.boxes):
                L.set_box(b, _field(n, bb[:3], bb[3:]))
    R = FluxRegister(C, F, 2)
    if init:
        C.average_down(F, 2)
        C.init_post()
        F.init_post()
    return C, F, R


def _steps(C, F, R, dt, n):
    for s in range(n):
        dt = min(C.initial_dt(), 2 * F.initial_dt()) if dt is None else \
            min(C.new_dt(dt), 2 * F.new_dt(dt / 2))
        amr_step(C, F, R, dt)
    return dt


def test_two_level_restart_is_bitwise(tmp_path):
    C, F, R = _hierarchy()
    dt = _steps(C, F, R, None, 2)
    path = str(tmp_path / "chk00002")
    write_checkpoint(path, [C, F], ref_ratio=[2], level_steps=[2, 4], dt_level=[dt, dt / 2])
    ck = read_checkpoint(path)
    assert ck["version"] == "CheckPointVersion_1.0" and ck["finest_level"] == 1
    assert ck["time"] == C.time and ck["ref_ratio"] == [2] and ck["level_steps"] == [2, 4]
    assert ck["dt_level"] == [dt, dt / 2]
    for rec, L in zip(ck["levels"], (C, F)):
        assert rec["time"] == L.time
        for b, a in enumerate(rec["data"]):
            assert np.array_equal(a, L.get_box(b))
    _steps(C, F, R, dt, 2)
    want = [C.get_box(b) for b in range(len(C.boxes))] + [F.get_box(b) for b in range(len(F.boxes))]
    C2, F2, R2 = _hierarchy(init=False)
    dts, steps = restart([C2, F2], path)
    assert C2.time == ck["time"] and steps == [2, 4]
    _steps(C2, F2, R2, dts[0], 2)
    got = [C2.get_box(b) for b in range(len(C2.boxes))] + [F2.get_box(b) for b in range(len(F2.boxes))]
    assert all(np.array_equal(a, b) for a, b in zip(want, got))
    assert C2.time == C.time
    for r in (R, R2):
        r.close()


def test_fused_level_restart_is_bitwise(tmp_path):
    prob = workloads.smooth_periodic(n=32, ppm_type=1, max_grid=16)
    A = Level(prob)
    for b in range(len(prob.boxes)):
        A.set_box(b, workloads.smooth_init(prob, b))
    A.init_post()
    dts = A.run(3)
    path = str(tmp_path / "chk")
    write_checkpoint(path, [A], dt_level=[dts[-1]])
    dt = dts[-1]
    for _ in range(2):
        dt = A.new_dt(dt)
        A.advance(dt)
    B = Level(prob)
    dt_level, _ = restart([B], path)
    dt = dt_level[0]
    for _ in range(2):
        dt = B.new_dt(dt)
        B.advance(dt)
    assert B.time == A.time
    for b in range(len(prob.boxes)):
        assert np.array_equal(B.get_box(b), A.get_box(b))
