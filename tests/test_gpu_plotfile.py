"""AMReX plotfiles of a two-level hierarchy (camr_write_plotfile, after
CAMR::writePlotFile, Source/Utils/CAMR_io.cpp:407-591) read back with the
package's reader: header fields, the level grids, state data bit for bit,
derived variables against their Derive.cpp formulas, per-box min/max."""
import re

import numpy as np
import pytest

from camr_amd import _lib
from camr_amd.hydro import FluxRegister, Level, Problem, amr_step
from camr_amd.plotfile import read_plotfile, write_plotfile
from tests.test_oracle_amr import FINE2, _boxes, _field

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

G = 1.4
STATE = ["density", "xmom", "ymom", "zmom", "rho_E", "rho_e", "Temp", "rho_F", "rho_A", "rho_P"]


def _hierarchy():
    def level(n, boxes):
        prob = Problem(domlo=(0, 0, 0), domhi=(n - 1,) * 3, dx=(1.0 / n,) * 3, bclo=(0,) * 3,
                       bchi=(0,) * 3, is_periodic=(1, 1, 1), boxes=boxes,
                       params=dict(ppm_type=1, cfl=0.3))
        L = Level(prob, engine=_lib.ENGINE_STAGED)
        for b, bb in enumerate(boxes):
            L.set_box(b, _field(n, bb[:3], bb[3:]))
        return L
    C, F = level(16, _boxes(16, 8)), level(32, FINE2)
    R = FluxRegister(C, F, 2)
    C.average_down(F, 2)
    C.init_post()
    F.init_post()
    dt = min(C.initial_dt(), 2 * F.initial_dt())
    amr_step(C, F, R, dt)
    return C, F, R


def _derived(U, name):
    rho, mx, my, mz, rhoe = U[0], U[1], U[2], U[3], U[5]
    if name.endswith("_velocity"):
        return U[1 + "xyz".index(name[0])] / rho
    sq = mx * mx + my * my + mz * mz
    if name == "kineng":
        return 0.5 * sq / rho
    p = (G - 1.0) * rho * (rhoe * (1.0 / rho))
    if name == "pressure":
        return p
    c = np.sqrt(G * p / rho)
    if name == "soundspeed":
        return c
    if name == "MachNumber":
        return np.sqrt(sq) / rho / c
    if name == "magvel":
        inv = 1.0 / rho
        return np.sqrt((mx * inv) ** 2 + (my * inv) ** 2 + (mz * inv) ** 2)
    raise KeyError(name)


def test_plotfile_round_trip(tmp_path):
    C, F, R = _hierarchy()
    path = tmp_path / "plt00001"
    ask = ["MachNumber", "x_velocity", "pressure", "kineng", "soundspeed", "y_velocity",
           "z_velocity", "magvel"]
    write_plotfile(str(path), [C, F], ref_ratio=[2], level_steps=[1, 2], plot_rhoy=True,
                   derive_vars=ask)
    pf = read_plotfile(str(path))
    order = ["pressure", "kineng", "soundspeed", "MachNumber", "x_velocity", "y_velocity",
             "z_velocity", "magvel"]  # CAMR's derive_lst order, not the request's
    assert pf["type"] == "HyperCLaw-V1.1"
    assert pf["names"] == STATE + order
    assert pf["time"] == C.time and pf["finest_level"] == 1
    assert pf["ref_ratio"] == [2] and pf["level_steps"] == [1, 2]
    assert pf["domains"] == [(0, 0, 0, 15, 15, 15), (0, 0, 0, 31, 31, 31)]
    assert pf["prob_hi"] == [1.0, 1.0, 1.0] and pf["dx"][1] == [1 / 32] * 3
    for lev, L in zip(pf["levels"], (C, F)):
        assert lev["time"] == L.time
        assert lev["boxes"] == [tuple(b) for b in L.boxes]
        dx = L.prob.dx[0]
        for b, bb in enumerate(L.boxes):
            loc = lev["grid_locations"][3 * b:3 * b + 3]
            assert loc == [[bb[d] * dx, (bb[3 + d] + 1) * dx] for d in range(3)]
            U = L.get_box(b)
            D = lev["data"][b]
            assert np.array_equal(D[:10], U)
            for k, name in enumerate(order):
                np.testing.assert_allclose(D[10 + k], _derived(U, name), rtol=1e-14, atol=0)
    # per-box min / max of the VisMF header
    vh = (path / "Level_1" / "Cell_H").read_text()
    blocks = vh.strip().split("\n\n")
    mins = [list(map(float, l.rstrip(",").split(","))) for l in blocks[-2].split("\n")[1:]]
    maxs = [list(map(float, l.rstrip(",").split(","))) for l in blocks[-1].split("\n")[1:]]
    for b, D in enumerate(pf["levels"][1]["data"]):
        assert mins[b] == [float(D[v].min()) for v in range(D.shape[0])]
        assert maxs[b] == [float(D[v].max()) for v in range(D.shape[0])]
    assert re.search(r"CAMR Job Information", (path / "job_info").read_text())
    R.close()


def test_plotfile_defaults_and_errors(tmp_path):
    C, F, R = _hierarchy()
    write_plotfile(str(tmp_path / "plt0"), [C])
    pf = read_plotfile(str(tmp_path / "plt0"))
    assert pf["names"] == STATE[:7] and pf["finest_level"] == 0
    assert np.array_equal(pf["levels"][0]["data"][3], C.get_box(3)[:7])
    with pytest.raises(_lib.CamrError, match="enstrophy"):
        write_plotfile(str(tmp_path / "plt1"), [C], derive_vars=["enstrophy"])
    R.close()
