"""The method-of-lines integrator (do_mol = 1) on the GPU against the
oracle, bit for bit: hydro_umdrv's MOL branch per box through the C ABI
(MOL_umeth.cpp:15-142 + Hydro_umdrv.cpp:65-92), and whole level steps with
the two-stage update of CAMR_advance.cpp:124-135 on the staged engine
(Sod tube, multi-box periodic, walls with a gravity source, the density
floor, captured-graph replays)."""
import ctypes as C

import numpy as np
import pytest

from camr_amd import _lib, workloads
from camr_amd.hydro import Level, Problem, hydro_params
from tests.test_gpu_parity import _box_state, _oracle_params, assert_bitwise

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _oparams(O, prob):
    p = _oracle_params(O, prob)
    p.do_mol = int(prob.value("do_mol"))
    return p


def _run_both(O, prob, init, nsteps, engine=_lib.ENGINE_AUTO, graph=False):
    ol = O.Level(_oparams(O, prob), prob.boxes)
    gl = Level(prob, engine=engine, graph=graph)
    for b in range(len(prob.boxes)):
        U = init(prob, b)
        ol.set_box(b, U)
        gl.set_box(b, U)
    ol.init_post()
    gl.init_post()
    err, odts = ol.run(nsteps)
    assert err == 0
    gdts = gl.run(nsteps)
    return ol, gl, odts, gdts


CASES = [
    dict(order=4, bclo=(0, 0, 0), bchi=(0, 0, 0)),
    dict(order=4, bclo=(4, 2, 5), bchi=(3, 4, 2)),
    dict(order=1, bclo=(5, 3, 2), bchi=(2, 5, 4)),
]


@pytest.mark.parametrize("case", CASES)
def test_umdrv_mol_per_box_bitwise(oracle, case):
    print(f"MOL per box {case}", flush=True)
    lo, hi = (3, 5, 7), (3 + 15, 5 + 11, 7 + 13)
    G = 4
    glo = tuple(l - G for l in lo)
    ghi = tuple(h + G for h in hi)
    U = _box_state(glo, ghi, seed=13)
    dx = (1 / 64,) * 3
    prob = Problem(domlo=lo, domhi=hi, dx=dx, bclo=case["bclo"], bchi=case["bchi"],
                   is_periodic=(0, 0, 0), boxes=[lo + hi],
                   params=dict(do_mol=1, plm_iorder=case["order"], cfl=0.3))
    hp = hydro_params(prob)
    assert hp.do_mol == 1
    op = _oparams(oracle, prob)
    dt = 0.3 * dx[0]
    q = np.zeros((11,) + U.shape[1:])
    qa = np.zeros((5,) + U.shape[1:])
    srcq = np.zeros_like(q)
    err = C.c_int(0)
    L3 = C.c_int * 3
    oracle.lib().ora_ctoprim_box(L3(*glo), L3(*ghi), C.byref(oracle.fab(U, glo)),
                                 C.byref(oracle.fab(q, glo)), C.byref(oracle.fab(qa, glo)),
                                 C.byref(op), C.byref(err))
    assert err.value == 0
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    dsdt = np.zeros((10, n[2], n[1], n[0]))
    flx = []
    for d in range(3):
        s = [n[2], n[1], n[0]]
        s[2 - d] += 1
        flx.append(np.zeros([10] + s))
    fl = (oracle.OraFab * 3)(*[oracle.fab(f, lo) for f in flx])
    oracle.lib().ora_hydro_umdrv(L3(*lo), L3(*hi), C.byref(oracle.fab(U, glo)),
                                 C.byref(oracle.fab(dsdt, lo)), C.byref(oracle.fab(q, glo)),
                                 C.byref(oracle.fab(qa, glo)), C.byref(oracle.fab(srcq, glo)),
                                 fl, dt, C.byref(op))
    from camr_amd.hydro import hydro_ctoprim, hydro_umdrv
    dev = torch.device("cuda:0")
    tU = torch.from_numpy(U).to(dev)
    tq = torch.zeros((11,) + U.shape[1:], dtype=torch.float64, device=dev)
    tqa = torch.zeros((5,) + U.shape[1:], dtype=torch.float64, device=dev)
    hydro_ctoprim(glo, ghi, tU, glo, tq, tqa, hp)
    tds = torch.zeros(dsdt.shape, dtype=torch.float64, device=dev)
    tfl = [torch.full(f.shape, float("nan"), dtype=torch.float64, device=dev) for f in flx]
    hydro_umdrv(lo, hi, tU, glo, tds, tq, tqa, None, tfl, dt, hp)
    torch.cuda.synchronize()
    for d in range(3):
        assert_bitwise(tfl[d].cpu().numpy(), flx[d], f"MOL flux {d}")
    assert_bitwise(tds.cpu().numpy(), dsdt, "MOL dsdt")
    assert np.abs(dsdt).max() > 0.0


def test_mol_sod_tube_bitwise(oracle):
    prob = workloads.sod_tube(cfl=0.3)
    prob.params["do_mol"] = 1
    print("MOL Sod tube 256x4x4, 60 steps", flush=True)
    ol, gl, odts, gdts = _run_both(oracle, prob, workloads.sod_init, 60)
    assert np.array_equal(odts, gdts)
    assert_bitwise(gl.get_box(0), ol.get_box(0), "MOL Sod state after 60 steps")


@pytest.mark.parametrize("graph", [False, True])
def test_mol_multibox_periodic_bitwise(oracle, graph):
    prob = workloads.smooth_periodic(n=32, ppm_type=0, cfl=0.3, max_grid=16)
    prob.params["do_mol"] = 1
    assert len(prob.boxes) == 8
    print(f"MOL smooth 32^3, 8 boxes, graph={graph}", flush=True)
    ol, gl, odts, gdts = _run_both(oracle, prob, workloads.smooth_init_varp, 5, graph=graph)
    assert np.array_equal(odts, gdts)
    for b in range(8):
        assert_bitwise(gl.get_box(b), ol.get_box(b), f"MOL box {b}")


@pytest.mark.parametrize("grav", [0.0, -0.7])
def test_mol_walls_and_sources_bitwise(oracle, grav):
    from camr_amd import problems
    n = (24, 20, 16)
    dx = (1 / 24,) * 3
    prob = Problem(domlo=(0, 0, 0), domhi=tuple(x - 1 for x in n), dx=dx,
                   bclo=(2, 4, 5), bchi=(3, 5, 4), is_periodic=(0, 0, 0),
                   boxes=[(0, 0, 0, 11, 19, 15), (12, 0, 0, 23, 19, 15)],
                   params=dict(do_mol=1, cfl=0.3, const_grav=grav))
    full = problems.SmoothField(seed=3, kmax=3, mach=0.4).state(48, (0, 0, 0), (23, 19, 15))

    def init(p, b):
        lo, hi = p.boxes[b][:3], p.boxes[b][3:]
        return np.ascontiguousarray(full[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1])

    print(f"MOL walls grav={grav}", flush=True)
    ol, gl, odts, gdts = _run_both(oracle, prob, init, 4)
    assert np.array_equal(odts, gdts)
    for b in range(2):
        assert_bitwise(gl.get_box(b), ol.get_box(b), f"MOL walls box {b}")


def test_mol_density_floor_path_bitwise(oracle):
    prob = workloads.smooth_periodic(n=16, ppm_type=0, cfl=0.3, max_grid=8)
    prob.params["do_mol"] = 1

    def init(p, b):
        U = workloads.smooth_init(p, b)
        if b == 0:
            for (i, j, k), rho in [((3, 3, 3), 5e-8), ((0, 0, 0), 2e-8), ((1, 0, 0), 2e-8)]:
                f = rho / U[0, k, j, i]
                U[[0, 1, 2, 3, 5, 7, 8, 9], k, j, i] *= f
                U[4, k, j, i] = U[5, k, j, i] + 0.5 * (U[1:4, k, j, i] ** 2).sum() / U[0, k, j, i]
        return U

    print("MOL density floor", flush=True)
    ol, gl, odts, gdts = _run_both(oracle, prob, init, 3)
    assert np.array_equal(odts, gdts)
    for b in range(len(prob.boxes)):
        assert_bitwise(gl.get_box(b), ol.get_box(b), f"MOL floor box {b}")


def test_mol_rejects_large_cfl_fused_engine_and_eb():
    prob = workloads.smooth_periodic(n=16, ppm_type=0, cfl=0.5, max_grid=16)
    prob.params["do_mol"] = 1
    with pytest.raises(_lib.CamrError, match="0.3"):
        Level(prob)
    prob.params["cfl"] = 0.3
    with pytest.raises(_lib.CamrError):
        Level(prob, engine=_lib.ENGINE_FUSED)
    sp = workloads.sod_sphere(n=16, max_grid=16)
    sp.params["do_mol"] = 1
    gl = Level(sp, engine=_lib.ENGINE_STAGED)
    with pytest.raises(_lib.CamrError, match="MOL"):
        gl.set_eb(workloads.sod_sphere_geometry(sp))
