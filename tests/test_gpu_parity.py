"""Parity of the gfx950 path with the CPU oracle, through the C-ABI.

Bar: bit-identical FP64 results (the path's stated tolerance is 1e-12 relative
L-inf per conserved field; the kernels reproduce the oracle's operation
order with FMA contraction disabled, so equality is exact).  Reported on
failure: max relative difference per component.
"""
import ctypes as C

import numpy as np
import pytest

from camr_amd import _lib, problems, workloads
from camr_amd.hydro import Level, Problem, hydro_params

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _rel(a, b):
    d = np.abs(a - b)
    s = np.maximum(np.abs(a), np.abs(b))
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.where(s > 0, d / s, 0.0)
    return r.reshape(r.shape[0], -1).max(axis=1)


def assert_bitwise(gpu, ref, what):
    if not np.array_equal(gpu, ref):
        raise AssertionError(f"{what}: not bit-identical; max rel diff per "
                             f"component {_rel(gpu, ref)}")


def _oracle_params(O, prob):
    return O.params(
        ppm_type=prob.value("ppm_type"), plm_iorder=prob.value("plm_iorder"),
        use_flattening=prob.value("use_flattening"), use_pslope=prob.value("use_pslope"),
        transverse_reset_density=prob.value("transverse_reset_density"),
        cfl=prob.value("cfl"), add_grav_src=prob.value("add_grav_src"),
        const_grav=prob.value("const_grav"), bclo=prob.bclo, bchi=prob.bchi,
        domlo=prob.domlo, domhi=prob.domhi, is_periodic=prob.is_periodic, dx=prob.dx)


def _run_both(O, prob, init, nsteps, engine=_lib.ENGINE_AUTO, batched=True, ext=None,
              tuning=None, graph=False):
    ol = O.Level(_oracle_params(O, prob), prob.boxes)
    gl = Level(prob, engine=engine, batched=batched, graph=graph)
    if tuning is not None:
        if len(tuning) > 2:
            from camr_amd import autotune
            autotune.apply(gl, *tuning)
        else:
            gl.tuning = tuning
    if ext is not None:
        ol.set_exterior(ext)
        gl.set_exterior(ext)
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


# ------------------------------------------------------------ per box
def _box_state(lo, hi, seed, shock=True):
    rng = np.random.default_rng(seed)
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    fld = problems.SmoothField(seed=seed, kmax=4, mach=0.6)
    N = 64
    U = fld.state(N, [l % N for l in lo], [l % N + n[d] - 1 for d, l in enumerate(lo)])
    if shock:
        # planar jumps in pressure and density to switch limiters / flattening
        z = np.arange(n[2])[:, None, None] + 0.5 * np.arange(n[0])[None, None, :]
        jump = (z > n[2] / 2).astype(float)
        rho = U[0] * (1.0 + 1.5 * jump)
        U[1:4] *= (rho / U[0])[None]
        U[0] = rho
        U[5] = U[5] * (1.0 + 4.0 * jump)
        U[7] = rho
    U[4] = U[5] + 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / U[0]
    U[6] = rng.uniform(0, 1, U[6].shape)  # UTEMP is overwritten by ctoprim
    return np.ascontiguousarray(U)


CASES = [
    dict(ppm=1, order=4, flat=1, bclo=(0, 0, 0), bchi=(0, 0, 0)),
    dict(ppm=0, order=4, flat=1, bclo=(0, 0, 0), bchi=(0, 0, 0)),
    dict(ppm=0, order=2, flat=0, bclo=(4, 2, 5), bchi=(3, 4, 2)),
    dict(ppm=1, order=4, flat=1, bclo=(5, 3, 2), bchi=(2, 5, 4)),
    # plm_pslope (PLM.H:84-165) with a non-zero primitive source
    dict(ppm=0, order=4, flat=1, pslope=1, src=True, bclo=(0, 0, 0), bchi=(0, 0, 0)),
    dict(ppm=0, order=2, flat=1, pslope=1, src=True, trd=0, bclo=(4, 2, 5), bchi=(3, 4, 2)),
    # transverse_reset_density = 0, sources through transdd's hdt * srcQ
    dict(ppm=1, order=4, flat=1, src=True, trd=0, bclo=(0, 0, 0), bchi=(0, 0, 0)),
    # a density band whose thin x-face states empty under the y flux
    # difference: the transdo density reset (Godunov_utils_3D.H:148-155)
    # fires
    dict(ppm=0, order=4, flat=1, pocket=True, trd=1, bclo=(0, 0, 0), bchi=(0, 0, 0)),
    dict(ppm=1, order=4, flat=1, pocket=True, trd=1, bclo=(0, 0, 0), bchi=(0, 0, 0)),
]


def _thinning_band(U):
    """Density falling tenfold per cell along x (to 1e-3), uniform pressure,
    in a flow that diverges along y: the limited slopes put a tenth of a
    cell's density on its thin x-face, and the y mass-flux difference of
    the cell exceeds that over one transverse correction."""
    nz, ny, nx = U.shape[1:]
    i, j = np.arange(nx), np.arange(ny)
    rho = np.clip(10.0 ** (-(i - 10.0)), 1e-3, 1.0)[None, None, :] * np.ones((nz, ny, 1))
    v = np.tanh((j - ny / 2 + 0.5) / 1.5)[None, :, None] * np.ones((nz, 1, nx))
    u, w = 0.2 * U[1] / U[0], 0.2 * U[3] / U[0]
    V = np.zeros_like(U)
    V[0], V[1], V[2], V[3] = rho, rho * u, rho * v, rho * w
    V[5] = 1.0 / 0.4
    V[7] = rho
    V[4] = V[5] + 0.5 * (V[1] ** 2 + V[2] ** 2 + V[3] ** 2) / V[0]
    V[6] = U[6]
    return np.ascontiguousarray(V)


def _oracle_umdrv(oracle, U, glo, ghi, lo, hi, op, dt, src):
    """ctoprim (+ srctoprim) and hydro_umdrv of the oracle on one box."""
    q = np.zeros((11,) + U.shape[1:])
    qa = np.zeros((5,) + U.shape[1:])
    srcq = np.zeros_like(q)
    err = C.c_int(0)
    L3 = C.c_int * 3
    oracle.lib().ora_ctoprim_box(L3(*glo), L3(*ghi), C.byref(oracle.fab(U, glo)),
                                 C.byref(oracle.fab(q, glo)), C.byref(oracle.fab(qa, glo)),
                                 C.byref(op), C.byref(err))
    assert err.value == 0
    if src is not None:
        oracle.lib().ora_srctoprim_box(L3(*glo), L3(*ghi), C.byref(oracle.fab(q, glo)),
                                       C.byref(oracle.fab(qa, glo)),
                                       C.byref(oracle.fab(src, glo)),
                                       C.byref(oracle.fab(srcq, glo)))
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
    return q, qa, srcq, dsdt, flx


@pytest.mark.parametrize("case", CASES)
def test_umdrv_per_box_bitwise(oracle, case):
    lo, hi = (3, 5, 7), (3 + 15, 5 + 11, 7 + 13)
    G = 4
    glo = tuple(l - G for l in lo)
    ghi = tuple(h + G for h in hi)
    U = _box_state(glo, ghi, seed=11)
    if case.get("pocket"):
        U = _thinning_band(U)
    src = None
    if case.get("src"):
        # a gravity-like conserved source (rho g along z, its work) plus a
        # smooth internal-energy source
        g = -0.7
        src = np.zeros_like(U)
        src[3] = U[0] * g
        src[4] = U[3] * g
        src[5] = 0.05 * np.sin(np.arange(U.shape[3]) / 3.0)[None, None, :] * U[0]
    dx = (1 / 64,) * 3
    prob = Problem(domlo=lo, domhi=hi, dx=dx, bclo=case["bclo"], bchi=case["bchi"],
                   is_periodic=(0, 0, 0), boxes=[lo + hi],
                   params=dict(ppm_type=case["ppm"], plm_iorder=case["order"],
                               use_flattening=case["flat"], use_pslope=case.get("pslope", 0),
                               transverse_reset_density=case.get("trd", 1)))
    hp = hydro_params(prob)
    op = _oracle_params(oracle, prob)
    dt = 0.3 * dx[0]
    q, qa, srcq, dsdt, flx = _oracle_umdrv(oracle, U, glo, ghi, lo, hi, op, dt, src)
    if case.get("pocket"):
        # the reset changes the result, so it fires on this state
        op0 = _oracle_params(oracle, prob)
        op0.transverse_reset_density = 0
        assert not np.array_equal(_oracle_umdrv(oracle, U, glo, ghi, lo, hi, op0, dt, src)[3],
                                  dsdt)
    # gpu
    from camr_amd.hydro import hydro_ctoprim, hydro_srctoprim, hydro_umdrv
    dev = torch.device("cuda:0")
    tU = torch.from_numpy(U).to(dev)
    tq = torch.zeros((11,) + U.shape[1:], dtype=torch.float64, device=dev)
    tqa = torch.zeros((5,) + U.shape[1:], dtype=torch.float64, device=dev)
    hydro_ctoprim(glo, ghi, tU, glo, tq, tqa, hp)
    torch.cuda.synchronize()
    assert_bitwise(tq.cpu().numpy(), q, "ctoprim q")
    assert_bitwise(tqa.cpu().numpy(), qa, "ctoprim qaux")
    tsq = None
    if src is not None:
        tsq = torch.zeros_like(tq)
        hydro_srctoprim(glo, ghi, tq, tqa, torch.from_numpy(src).to(dev), tsq)
        torch.cuda.synchronize()
        assert_bitwise(tsq.cpu().numpy(), srcq, "srctoprim")
    tds = torch.zeros(dsdt.shape, dtype=torch.float64, device=dev)
    tfl = [torch.full(f.shape, float("nan"), dtype=torch.float64, device=dev) for f in flx]
    hydro_umdrv(lo, hi, tU, glo, tds, tq, tqa, tsq, tfl, dt, hp)
    torch.cuda.synchronize()
    for d in range(3):
        assert_bitwise(tfl[d].cpu().numpy(), flx[d], f"flux {d}")
    assert_bitwise(tds.cpu().numpy(), dsdt, "dsdt")


@pytest.mark.parametrize("route", ["area_vol_arrays", "larger_q_box"])
def test_umdrv_stage_route_bitwise(oracle, route):
    """Calls that leave the fused route (area / volume arrays given, or q
    over more than grow(bx,4)) run the stage-by-stage kernels, with the
    same results."""
    lo, hi = (3, 5, 7), (3 + 15, 5 + 11, 7 + 13)
    G = 5 if route == "larger_q_box" else 4
    glo = tuple(l - G for l in lo)
    ghi = tuple(h + G for h in hi)
    U = _box_state(glo, ghi, seed=13)
    dx = (1 / 64,) * 3
    prob = Problem(domlo=lo, domhi=hi, dx=dx, bclo=(0, 0, 0), bchi=(0, 0, 0),
                   is_periodic=(0, 0, 0), boxes=[lo + hi], params=dict(ppm_type=1))
    hp = hydro_params(prob)
    op = _oracle_params(oracle, prob)
    dt = 0.3 * dx[0]
    q, qa, srcq, dsdt, flx = _oracle_umdrv(oracle, U, glo, ghi, lo, hi, op, dt, None)
    from camr_amd.hydro import hydro_ctoprim, hydro_umdrv
    dev = torch.device("cuda:0")
    tU = torch.from_numpy(U).to(dev)
    tq = torch.zeros((11,) + U.shape[1:], dtype=torch.float64, device=dev)
    tqa = torch.zeros((5,) + U.shape[1:], dtype=torch.float64, device=dev)
    hydro_ctoprim(glo, ghi, tU, glo, tq, tqa, hp)
    area = vol = None
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    if route == "area_vol_arrays":
        area = []
        for d in range(3):
            sh = [1, n[2], n[1], n[0]]
            sh[3 - d] += 1
            a = dx[(d + 1) % 3] * dx[(d + 2) % 3] if d != 1 else dx[0] * dx[2]
            area.append(torch.full(sh, a, dtype=torch.float64, device=dev))
        vol = torch.full([1, n[2], n[1], n[0]], dx[0] * dx[1] * dx[2], dtype=torch.float64,
                         device=dev)
    tds = torch.zeros(dsdt.shape, dtype=torch.float64, device=dev)
    tfl = [torch.full(f.shape, float("nan"), dtype=torch.float64, device=dev) for f in flx]
    # q and qaux start at lo - 4 for the call: pass the grow(bx,4) window
    # unless the route under test is the larger box
    if route == "larger_q_box":
        from camr_amd import _lib as LL
        from camr_amd.hydro import array4_of, _box, _stream
        import ctypes as CC
        fl = (LL.Array4 * 3)(*[array4_of(f, lo) for f in tfl])
        ar = (LL.Array4 * 3)(*[LL.Array4() for _ in range(3)])
        LL.check(LL.lib().camr_hydro_umdrv(
            CC.byref(_box(lo, hi)), CC.byref(array4_of(tU, glo)), CC.byref(array4_of(tds, lo)),
            CC.byref(array4_of(tq, glo)), CC.byref(array4_of(tqa, glo)), CC.byref(LL.Array4()),
            fl, ar, CC.byref(LL.Array4()), dt, CC.byref(hp), _stream(0)))
    else:
        hydro_umdrv(lo, hi, tU, glo, tds, tq, tqa, None, tfl, dt, hp, area=area, vol=vol)
    torch.cuda.synchronize()
    for d in range(3):
        assert_bitwise(tfl[d].cpu().numpy(), flx[d], f"flux {d}")
    assert_bitwise(tds.cpu().numpy(), dsdt, "dsdt")


def test_estdt_matches_oracle(oracle):
    lo, hi = (0, 0, 0), (20, 17, 9)
    U = _box_state(lo, hi, seed=5)
    dx = (0.01, 0.01, 0.01)
    from camr_amd.hydro import estdt_hydro
    t = torch.from_numpy(U).to("cuda:0")
    g = estdt_hydro(lo, hi, t, lo, dx)
    L3 = C.c_int * 3
    o = oracle.lib().ora_estdt_box(L3(*lo), L3(*hi), C.byref(oracle.fab(U, lo)),
                                   (C.c_double * 3)(*dx))
    assert g == o


# ------------------------------------------------------------ level steps
@pytest.mark.parametrize("engine", [_lib.ENGINE_AUTO, _lib.ENGINE_STAGED])
def test_sod_tube_100_steps_bitwise(oracle, engine):
    prob = workloads.sod_tube()
    ol, gl, odts, gdts = _run_both(oracle, prob, workloads.sod_init, 100, engine)
    assert np.array_equal(odts, gdts), (odts - gdts)
    assert_bitwise(gl.get_box(0), ol.get_box(0), "Sod state after 100 steps")


@pytest.mark.parametrize("ppm", [1, 0])
def test_smooth_periodic_bitwise(oracle, ppm):
    prob = workloads.smooth_periodic(n=32, ppm_type=ppm)
    ol, gl, odts, gdts = _run_both(oracle, prob, workloads.smooth_init, 5)
    assert np.array_equal(odts, gdts)
    assert_bitwise(gl.get_box(0), ol.get_box(0), "smooth 32^3 after 5 steps")


@pytest.mark.parametrize("batched,graph", [(True, False), (False, False), (True, True)])
def test_multibox_level_bitwise(oracle, batched, graph):
    prob = workloads.smooth_periodic(n=32, ppm_type=1, max_grid=16)
    assert len(prob.boxes) == 8
    ol, gl, odts, gdts = _run_both(oracle, prob, workloads.smooth_init, 5, batched=batched,
                                   graph=graph)
    assert np.array_equal(odts, gdts)
    for b in range(8):
        assert_bitwise(gl.get_box(b), ol.get_box(b), f"box {b}")


@pytest.mark.parametrize("engine", [_lib.ENGINE_AUTO, _lib.ENGINE_STAGED])
@pytest.mark.parametrize("grav,pslope,trd", [(0.0, 0, 1), (-0.7, 0, 1), (-0.7, 1, 1),
                                             (-0.7, 1, 0)])
def test_walls_and_sources_bitwise(oracle, engine, grav, pslope, trd):
    n = (24, 20, 16)
    dx = (1 / 24,) * 3
    prob = Problem(domlo=(0, 0, 0), domhi=tuple(x - 1 for x in n), dx=dx,
                   bclo=(2, 4, 5), bchi=(3, 5, 4), is_periodic=(0, 0, 0),
                   boxes=[(0, 0, 0, 11, 19, 15), (12, 0, 0, 23, 19, 15)],
                   params=dict(ppm_type=0, cfl=0.4, const_grav=grav, use_pslope=pslope,
                               transverse_reset_density=trd))
    full = problems.SmoothField(seed=3, kmax=3, mach=0.4).state(48, (0, 0, 0), (23, 19, 15))

    def init(p, b):
        lo, hi = p.boxes[b][:3], p.boxes[b][3:]
        return np.ascontiguousarray(full[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1])

    ol, gl, odts, gdts = _run_both(oracle, prob, init, 4, engine)
    assert np.array_equal(odts, gdts)
    for b in range(2):
        assert_bitwise(gl.get_box(b), ol.get_box(b), f"walls box {b}")


def test_large_periodic_conservation():
    # size-independent property at the full 256^3 size: periodic fluxes
    # telescope, so mass, momentum and total energy are conserved to
    # round-off; species 0 tracks density.
    prob = workloads.smooth_periodic(n=256, ppm_type=1)
    gl = Level(prob)
    U = workloads.smooth_init(prob, 0)
    gl.set_box(0, U)
    gl.init_post()
    s0 = gl.get_box(0).reshape(10, -1).sum(axis=1)
    gl.run(2)
    S = gl.get_box(0)
    s1 = S.reshape(10, -1).sum(axis=1)
    scale = np.abs(S).reshape(10, -1).sum(axis=1)
    for c in (0, 1, 2, 3, 4):
        assert abs(s1[c] - s0[c]) <= 1e-12 * scale[c], (c, s1[c] - s0[c])
    assert np.all(np.isfinite(S))


@pytest.mark.parametrize("engine,graph", [(_lib.ENGINE_AUTO, False), (_lib.ENGINE_STAGED, False),
                                          (_lib.ENGINE_AUTO, True)])
def test_density_floor_path_bitwise(oracle, engine, graph):
    # cells below small_dens (1e-7) trigger enforce_min_density's neighbour
    # search / small-state reset; the fused engine detects them and redoes
    # the step with the exact pass-by-pass sequence
    prob = workloads.smooth_periodic(n=16, ppm_type=0, max_grid=8)

    def init(p, b):
        U = workloads.smooth_init(p, b)
        if b == 0:
            for (i, j, k), rho in [((3, 3, 3), 5e-8), ((0, 0, 0), 2e-8), ((1, 0, 0), 2e-8)]:
                f = rho / U[0, k, j, i]
                U[[0, 1, 2, 3, 5, 7, 8, 9], k, j, i] *= f
                U[4, k, j, i] = U[5, k, j, i] + 0.5 * (U[1:4, k, j, i] ** 2).sum() / U[0, k, j, i]
        return U

    ol, gl, odts, gdts = _run_both(oracle, prob, init, 3, engine, graph=graph)
    assert np.array_equal(odts, gdts)
    for b in range(len(prob.boxes)):
        assert_bitwise(gl.get_box(b), ol.get_box(b), f"floor box {b}")


# ------------------------------------------------- batched refined level
@pytest.mark.parametrize("batched,tuning,graph", [
    (True, None, False),
    (False, None, False),
    (True, ([64, 128, 256, 64, 128, 256, 64, 128, 64, 256, 128, 64, 128], 1), False),
    (True, None, True),
])
def test_refined_level_with_exterior_bitwise(oracle, batched, tuning, graph):
    # uneven boxes (sides 8..32) covering a sphere of a 64^3 periodic index
    # space; ghost cells outside the level come from exterior data
    prob = workloads.refined_level(n=64, radius=0.3, blocking_factor=8, max_size=32,
                                   efficiency=0.8)
    assert len(prob.boxes) > 10
    ext = workloads.exterior(prob)

    def init(p, b):
        lo, hi = p.boxes[b][:3], p.boxes[b][3:]
        return np.ascontiguousarray(ext[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1])

    ol, gl, odts, gdts = _run_both(oracle, prob, init, 3, batched=batched, ext=ext,
                                   tuning=tuning, graph=graph)
    assert np.array_equal(odts, gdts)
    for b in range(len(prob.boxes)):
        assert_bitwise(gl.get_box(b), ol.get_box(b), f"refined box {b}")


TILE_SETS = {
    "16x4x4": [(16, 4, 4)] * 13,
    "mixed": [(64, 4, 1), (32, 4, 2), (16, 8, 2), (8, 8, 4), (32, 2, 4), (16, 4, 4),
              (0, 0, 0), (32, 8, 1), (64, 2, 2), (16, 4, 2), (32, 2, 2), (8, 8, 2), (32, 4, 1)],
}


@pytest.mark.parametrize("tiles", sorted(TILE_SETS))
@pytest.mark.parametrize("batched,graph", [(True, False), (False, False), (True, True)])
def test_tiled_blocks_bitwise(oracle, tiles, batched, graph):
    """3-D tiled blocks (the autotuner's candidates, the last partial tiles
    of uneven boxes included) change which thread computes a point, not
    its value: a refined level of uneven boxes, bit for bit."""
    prob = workloads.refined_level(n=64, radius=0.3, blocking_factor=8, max_size=32,
                                   efficiency=0.8)
    ext = workloads.exterior(prob)

    def init(p, b):
        lo, hi = p.boxes[b][:3], p.boxes[b][3:]
        return np.ascontiguousarray(ext[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1])

    ol = oracle.Level(_oracle_params(oracle, prob), prob.boxes)
    gl = Level(prob, batched=batched, graph=graph)
    gl.tiles = TILE_SETS[tiles]
    assert gl.tiles == TILE_SETS[tiles]
    if graph:
        gl.ctu = (0, 32)              # tiles of the staged transverse launches too
    # block permutations over the XCDs, chunks of 1 to 64 blocks
    gl.xcd = [0, 1, 2, 3, 4, 8, 16, 32, 64, 5, 7, 0, 12]
    for L in (ol, gl):
        L.set_exterior(ext)
        for b in range(len(prob.boxes)):
            L.set_box(b, init(prob, b))
        L.init_post()
    err, odts = ol.run(3)
    assert err == 0
    gdts = gl.run(3)
    assert np.array_equal(odts, gdts)
    for b in range(len(prob.boxes)):
        assert_bitwise(gl.get_box(b), ol.get_box(b), f"tiled box {b}")


@pytest.mark.parametrize("planes", [1, 3, 8, 64])
@pytest.mark.parametrize("batched,graph", [(True, False), (False, False), (True, True)])
def test_zmarch_bitwise(oracle, planes, batched, graph):
    """z-face transverse stages with every thread walking `planes` z planes
    (the last runs of uneven boxes shorter), values carried between faces in
    registers: a refined level of uneven boxes, bit for bit."""
    prob = workloads.refined_level(n=64, radius=0.3, blocking_factor=8, max_size=32,
                                   efficiency=0.8)
    ext = workloads.exterior(prob)

    def init(p, b):
        lo, hi = p.boxes[b][:3], p.boxes[b][3:]
        return np.ascontiguousarray(ext[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1])

    ol = oracle.Level(_oracle_params(oracle, prob), prob.boxes)
    gl = Level(prob, batched=batched, graph=graph)
    zm = [0] * 13
    zm[7] = zm[11] = planes
    gl.ctu = (0, 32)                  # the staged transverse launches
    gl.zmarch = zm
    gl.xcd = [16] * 13
    assert gl.zmarch == zm
    for L in (ol, gl):
        L.set_exterior(ext)
        for b in range(len(prob.boxes)):
            L.set_box(b, init(prob, b))
        L.init_post()
    err, odts = ol.run(3)
    assert err == 0
    gdts = gl.run(3)
    assert np.array_equal(odts, gdts)
    for b in range(len(prob.boxes)):
        assert_bitwise(gl.get_box(b), ol.get_box(b), f"z-marched box {b}")


@pytest.mark.parametrize("ppm", [0, 1])
def test_zmarch_with_sources_and_walls_bitwise(oracle, ppm):
    n = (24, 20, 16)
    dx = (1 / 24,) * 3
    prob = Problem(domlo=(0, 0, 0), domhi=tuple(x - 1 for x in n), dx=dx,
                   bclo=(2, 4, 5), bchi=(3, 5, 4), is_periodic=(0, 0, 0),
                   boxes=[(0, 0, 0, 11, 19, 15), (12, 0, 0, 23, 19, 15)],
                   params=dict(ppm_type=ppm, cfl=0.4, const_grav=-0.7))
    full = problems.SmoothField(seed=3, kmax=3, mach=0.4).state(48, (0, 0, 0), (23, 19, 15))
    ol = oracle.Level(_oracle_params(oracle, prob), prob.boxes)
    gl = Level(prob, engine=_lib.ENGINE_FUSED)
    gl.ctu = (0, 32)
    zm = [0] * 13
    zm[7] = zm[11] = 5
    gl.zmarch = zm
    for L in (ol, gl):
        for b in range(2):
            lo, hi = prob.boxes[b][:3], prob.boxes[b][3:]
            L.set_box(b, np.ascontiguousarray(full[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1,
                                                   lo[0]:hi[0] + 1]))
        L.init_post()
    err, odts = ol.run(4)
    assert err == 0
    gdts = gl.run(4)
    assert np.array_equal(odts, gdts)
    for b in range(2):
        assert_bitwise(gl.get_box(b), ol.get_box(b), f"z-marched walls box {b}")


@pytest.mark.parametrize("planes", [1, 5, 32, 300])
@pytest.mark.parametrize("batched,graph", [(True, False), (False, False), (True, True)])
def test_ctu_kernels_bitwise(oracle, planes, batched, graph):
    """The LDS-staged CTU face kernels (transdo pairs, transdd and the final
    Riemann solve of a face direction in one kernel) with blocks walking
    `planes` z planes: a refined level of uneven boxes (partial tiles, merged
    x tiles, boxes shorter than a block's run), bit for bit."""
    prob = workloads.refined_level(n=64, radius=0.3, blocking_factor=8, max_size=32,
                                   efficiency=0.8)
    ext = workloads.exterior(prob)

    def init(p, b):
        lo, hi = p.boxes[b][:3], p.boxes[b][3:]
        return np.ascontiguousarray(ext[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1])

    ol = oracle.Level(_oracle_params(oracle, prob), prob.boxes)
    gl = Level(prob, batched=batched, graph=graph)
    gl.ctu = (1, planes)
    gl.xcd = [8] * 13
    assert gl.ctu == (1, planes)
    for L in (ol, gl):
        L.set_exterior(ext)
        for b in range(len(prob.boxes)):
            L.set_box(b, init(prob, b))
        L.init_post()
    err, odts = ol.run(3)
    assert err == 0
    gdts = gl.run(3)
    assert np.array_equal(odts, gdts)
    for b in range(len(prob.boxes)):
        assert_bitwise(gl.get_box(b), ol.get_box(b), f"CTU box {b}")


@pytest.mark.parametrize("ppm,ctu", [(0, 1), (1, 1), (1, 0)])
def test_ctu_with_sources_and_walls_bitwise(oracle, ppm, ctu):
    """Walls, outflow and gravity (the transdd source terms) through the CTU
    face kernels and, for comparison, the staged launches; boxes of 40 x 33
    x 17 points leave partial tiles in every direction."""
    n = (80, 33, 17)
    dx = (1 / 80,) * 3
    prob = Problem(domlo=(0, 0, 0), domhi=tuple(x - 1 for x in n), dx=dx,
                   bclo=(2, 4, 5), bchi=(3, 5, 4), is_periodic=(0, 0, 0),
                   boxes=[(0, 0, 0, 39, 32, 16), (40, 0, 0, 79, 32, 16)],
                   params=dict(ppm_type=ppm, cfl=0.4, const_grav=-0.7))
    full = problems.SmoothField(seed=5, kmax=3, mach=0.4).state(80, (0, 0, 0), (79, 32, 16))
    ol = oracle.Level(_oracle_params(oracle, prob), prob.boxes)
    gl = Level(prob, engine=_lib.ENGINE_FUSED)
    gl.ctu = (ctu, 7)
    for L in (ol, gl):
        for b in range(2):
            lo, hi = prob.boxes[b][:3], prob.boxes[b][3:]
            L.set_box(b, np.ascontiguousarray(full[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1,
                                                   lo[0]:hi[0] + 1]))
        L.init_post()
    err, odts = ol.run(4)
    assert err == 0
    gdts = gl.run(4)
    assert np.array_equal(odts, gdts)
    for b in range(2):
        assert_bitwise(gl.get_box(b), ol.get_box(b), f"CTU walls box {b}")


def test_ctu_settings_are_validated():
    gl = Level(workloads.smooth_periodic(n=16, ppm_type=1))
    assert gl.ctu == (0, 8)
    for bad in ((2, 32), (1, 0), (1, 1025)):
        with pytest.raises(_lib.CamrError):
            gl.ctu = bad


def test_tile_shapes_are_validated():
    prob = workloads.smooth_periodic(n=16, ppm_type=1)
    gl = Level(prob)
    for bad in ([(12, 4, 4)] * 13, [(4, 8, 8)] * 13, [(64, 4, 4)] * 13):
        with pytest.raises(_lib.CamrError):
            gl.tiles = bad
    with pytest.raises(_lib.CamrError):
        gl.zmarch = [4] * 13          # only the z-face transverse stages march
    gl.tiles = [(16, 4, 4)] * 13
    zm = [0] * 13
    zm[7] = 4
    with pytest.raises(_lib.CamrError):
        gl.zmarch = zm                # marching needs linear blocks


def test_unowned_ghosts_without_exterior_are_an_error():
    prob = workloads.refined_level(n=32, radius=0.3, blocking_factor=8, max_size=16)
    gl = Level(prob)
    for b in range(len(prob.boxes)):
        lo, hi = prob.boxes[b][:3], prob.boxes[b][3:]
        gl.set_box(b, workloads.smooth_field().state(32, lo, hi))
    gl.init_post()
    with pytest.raises(_lib.CamrError):
        gl.run(1)


def test_graph_replays_continue_after_host_steps(oracle):
    # host-stepped and graph-replayed runs interleave on one level: the
    # device dt sequence picks up the host's, and back
    prob = workloads.smooth_periodic(n=24, ppm_type=1, max_grid=12)
    ol = oracle.Level(_oracle_params(oracle, prob), prob.boxes)
    gh = Level(prob, graph=False)
    gg = Level(prob, graph=True)
    for b in range(len(prob.boxes)):
        U = workloads.smooth_init(prob, b)
        ol.set_box(b, U)
        gh.set_box(b, U)
        gg.set_box(b, U)
    for L in (ol, gh, gg):
        L.init_post()
    err, odts = ol.run(6)
    assert err == 0
    hdts = gh.run(6)
    g1 = gg.run(4)
    dt = gg.new_dt(g1[-1])
    gg.advance(dt)
    g2 = [dt, gg.new_dt(dt)]
    gg.advance(g2[1])
    assert np.array_equal(odts, hdts)
    assert np.array_equal(odts, np.concatenate([g1, g2]))
    assert gg.time == ol.time
    for b in range(len(prob.boxes)):
        assert_bitwise(gg.get_box(b), ol.get_box(b), f"graph box {b}")


def test_autotuned_level_stays_bitwise(oracle, tmp_path, monkeypatch):
    from camr_amd import autotune
    monkeypatch.setenv("CAMR_TUNING_CACHE", str(tmp_path / "tuning.json"))
    prob = workloads.refined_level(n=32, radius=0.3, blocking_factor=8, max_size=16)
    ext = workloads.exterior(prob)

    def init(p, b):
        lo, hi = p.boxes[b][:3], p.boxes[b][3:]
        return np.ascontiguousarray(ext[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1])

    T = Level(prob)
    T.set_exterior(ext)
    for b in range(len(prob.boxes)):
        T.set_box(b, init(prob, b))
    T.init_post()
    res = autotune.tune(T, steps=1)
    blocks, order, tiles, xcd, ctu = res
    assert len(blocks) == _lib.NSTAGE and all(x in (64, 128, 256) for x in blocks)
    assert all(t == (0, 0, 0) or t[0] * t[1] * t[2] == b for b, t in zip(blocks, tiles))
    assert len(xcd) == _lib.NSTAGE
    assert ctu[0] in (0, 1) and ctu[1] >= 1
    assert (tmp_path / "tuning.json").exists()
    # a second level with the same boxes reads the cache
    assert autotune.tune(Level(prob), steps=1) == res
    ol, gl, odts, gdts = _run_both(oracle, prob, init, 3, ext=ext, tuning=res, graph=True)
    assert np.array_equal(odts, gdts)
    for b in range(len(prob.boxes)):
        assert_bitwise(gl.get_box(b), ol.get_box(b), f"tuned box {b}")


@pytest.mark.parametrize("graph,mol,grav", [(False, 0, 0.0), (True, 0, 0.0), (False, 1, 0.0),
                                            (True, 1, 0.0), (False, 0, -0.7)])
def test_staged_tail_floors_in_new_state_bitwise(oracle, graph, mol, grav):
    # small_dens inside the density range of the field: enforce_min_density
    # acts on S_new after the update every step, where the staged engine's
    # one-pass tail (with mol: the second stage of the method of lines)
    # hands over to its pass-by-pass repeat
    prob = workloads.smooth_periodic(n=16, ppm_type=0, max_grid=8)
    # (the method of lines at its CFL limit 0.3 moves the density less per
    # step: a floor at the field's mean density acts after its update)
    sd = 1.0 if mol else 0.95
    prob.params["small_dens"] = sd
    prob.params["do_mol"] = mol
    prob.params["const_grav"] = grav  # with gravity: the new-time source and its clean
    if mol:
        prob.params["cfl"] = 0.3  # the MOL integrator's limit
    p = _oracle_params(oracle, prob)
    p.small_dens = sd
    p.do_mol = mol
    ol = oracle.Level(p, prob.boxes)
    gl = Level(prob, engine=_lib.ENGINE_STAGED, graph=graph)
    for b in range(len(prob.boxes)):
        U = workloads.smooth_init(prob, b)
        ol.set_box(b, U)
        gl.set_box(b, U)
    ol.init_post()
    gl.init_post()
    err, odts = ol.run(3)
    assert err == 0
    gdts = gl.run(3)
    assert np.array_equal(odts, np.asarray(gdts))
    copied = 0
    for b in range(len(prob.boxes)):
        S = gl.get_box(b)
        assert_bitwise(S, ol.get_box(b), f"floor tail box {b}")
        # a floor copies a neighbour's whole state: equal neighbours (of the
        # 26) in the final S_new show that floors acted after the update
        n = S.shape[1:]
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    if dx == dy == dz == 0:
                        continue
                    a = S[:, max(dz, 0):n[0] + min(dz, 0), max(dy, 0):n[1] + min(dy, 0),
                          max(dx, 0):n[2] + min(dx, 0)]
                    c = S[:, max(-dz, 0):n[0] + min(-dz, 0), max(-dy, 0):n[1] + min(-dy, 0),
                          max(-dx, 0):n[2] + min(-dx, 0)]
                    copied += int(np.all(a == c, axis=0).sum())
    # (with gravity the new-time source, which differs cell by cell, acts
    # on the copies after the first floor pass, so they need not survive)
    assert copied > 0 or grav != 0.0
