"""Embedded boundaries on the GPU against the oracle, bit for bit.

The Sod-past-a-sphere problem (Exec/SodPlusSphere: Outflow in x, SlipWall
in y/z, PLM, FluxRedist, eb_weights_type 2) on small grids, with box layouts
that put cut cells in several boxes and next to box edges: every step's dt
and the final state of every box (covered cells included) must equal the
oracle's exactly.  The interface starts close to the sphere so that the
shock crosses the cut cells within the steps run.
"""
import numpy as np
import pytest

from camr_amd import _lib, workloads
from camr_amd.hydro import Level

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

UNTILED = (1 << 20,) * 3


def _oracle_params(O, prob):
    return O.params(
        ppm_type=prob.value("ppm_type"), plm_iorder=prob.value("plm_iorder"),
        use_flattening=prob.value("use_flattening"), use_pslope=prob.value("use_pslope"),
        transverse_reset_density=prob.value("transverse_reset_density"),
        cfl=prob.value("cfl"), bclo=prob.bclo, bchi=prob.bchi, domlo=prob.domlo,
        domhi=prob.domhi, is_periodic=prob.is_periodic, dx=prob.dx, tile=UNTILED)


@pytest.mark.parametrize("n,max_grid,nsteps", [(16, 16, 6), (32, 16, 5), (32, 8, 4), (24, 16, 4)])
def test_eb_level_matches_oracle(oracle, n, max_grid, nsteps):
    print(f"EB level n={n} max_grid={max_grid} steps={nsteps}", flush=True)
    prob = workloads.sod_sphere(n=n, max_grid=max_grid, interface=0.4)
    geom = workloads.sod_sphere_geometry(prob)
    ol = oracle.Level(_oracle_params(oracle, prob), prob.boxes)
    ol.set_eb(geom)
    gl = Level(prob, engine=_lib.ENGINE_STAGED)
    gl.set_eb(geom)
    assert gl.kinds == ol.kinds and 1 in gl.kinds
    for b in range(len(prob.boxes)):
        U = workloads.sod_sphere_init(prob, b, geom)
        ol.set_box(b, U)
        gl.set_box(b, U)
    ol.init_post()
    gl.init_post()
    err, odts = ol.run(nsteps)
    assert err == 0
    gdts = gl.run(nsteps)
    assert np.array_equal(np.asarray(gdts), odts)
    worst = 0.0
    for b in range(len(prob.boxes)):
        a, o = gl.get_box(b), ol.get_box(b)
        if not np.array_equal(a, o):
            d = np.abs(a - o).max() / max(np.abs(o).max(), 1e-300)
            worst = max(worst, d)
            print(f"box {b} kind {gl.kinds[b]} differs: max rel {d:.3e}", flush=True)
    assert worst == 0.0
    # the force on the sphere from the final state (the wall momentum flux
    # of every cut cell times dx^2), bit for bit
    gf, of = gl.eb_force(), ol.eb_force()
    assert np.array_equal(gf, of), (gf, of)
    assert abs(gf[0]) > 0.0
    # the shock reached the cut cells
    init = [workloads.sod_sphere_init(prob, b, geom) for b in range(len(prob.boxes))]
    moved = max(float(np.abs(gl.get_box(b)[0] - init[b][0]).max())
                for b in range(len(prob.boxes)) if gl.kinds[b] == 1)
    assert moved > 1e-3


def test_eb_level_graph_replays_match_oracle(oracle):
    """The EB level step captured in a graph (one per buffer parity) and
    replayed step after step, with host syncs and dt computations between
    the replays: every step's state equals the oracle's."""
    prob = workloads.sod_sphere(n=32, max_grid=16, interface=0.4)
    geom = workloads.sod_sphere_geometry(prob)
    ol = oracle.Level(_oracle_params(oracle, prob), prob.boxes)
    ol.set_eb(geom)
    gl = Level(prob, engine=_lib.ENGINE_STAGED, graph=True)
    gl.set_eb(geom)
    for b in range(len(prob.boxes)):
        U = workloads.sod_sphere_init(prob, b, geom)
        ol.set_box(b, U)
        gl.set_box(b, U)
    ol.init_post()
    gl.init_post()
    odt = gdt = None
    for s in range(5):
        odt = ol.initial_dt() if s == 0 else ol.new_dt(odt)
        gdt = gl.initial_dt() if s == 0 else gl.new_dt(gdt)
        assert odt == gdt
        assert ol.advance(odt) == 0
        gl.advance(gdt)
        for b in range(len(prob.boxes)):
            assert np.array_equal(gl.get_box(b), ol.get_box(b)), (s, b)


def test_eb_levels_are_refused_by_coarse_fine_transfers():
    """average_down and the coarse-fine ghost interpolation of EB builds
    (EB_average_down, eb_cell_cons_interp) are not built: the plain ones
    must refuse EB levels instead of averaging covered cells in."""
    crse = workloads.sod_sphere(n=16, max_grid=16)
    fine = workloads.sod_sphere(n=32, max_grid=16)
    C = Level(crse, engine=_lib.ENGINE_STAGED)
    C.set_eb(workloads.sod_sphere_geometry(crse))
    F = Level(fine, engine=_lib.ENGINE_STAGED)
    F.set_eb(workloads.sod_sphere_geometry(fine))
    with pytest.raises(_lib.CamrError):
        C.average_down(F, 2)
    with pytest.raises(_lib.CamrError):
        F.exterior_from_coarse(C, 2)


def test_eb_level_rejects_fused_engine_ppm_and_gravity():
    prob = workloads.sod_sphere(n=16, max_grid=16)
    geom = workloads.sod_sphere_geometry(prob)
    gl = Level(prob, engine=_lib.ENGINE_FUSED)
    with pytest.raises(_lib.CamrError):
        gl.set_eb(geom)
    prob.params["ppm_type"] = 1
    gl = Level(prob, engine=_lib.ENGINE_STAGED)
    with pytest.raises(_lib.CamrError):
        gl.set_eb(geom)
    # a constant-gravity source would need the cut-cell path's 6 ghost cells
    prob.params["ppm_type"] = 0
    prob.params["const_grav"] = -0.5
    gl = Level(prob, engine=_lib.ENGINE_STAGED)
    with pytest.raises(_lib.CamrError, match="gravity"):
        gl.set_eb(geom)


def test_umdrv_eb_per_box_bitwise(oracle):
    """camr_hydro_umdrv_eb through the C ABI against the oracle's
    hydro_umdrv_eb on one cut-cell box of Sod past a sphere, with a state
    that varies in all three directions (shock plus a smooth field)."""
    import ctypes as C
    from camr_amd import eb
    from camr_amd.hydro import hydro_ctoprim, hydro_params, hydro_umdrv_eb
    prob = workloads.sod_sphere(n=16, max_grid=16, interface=0.4)
    geom = workloads.sod_sphere_geometry(prob)
    b = 0
    lo, hi = prob.boxes[b][:3], prob.boxes[b][3:]
    g6lo = tuple(x - 6 for x in lo)
    g6hi = tuple(x + 6 for x in hi)
    assert geom.box_kind(b) == eb.BOX_EB
    # state on grow(bx, 6): Sod split in x + a smooth perturbation, covered cells zero
    x, y, z = [(np.arange(g6lo[d], g6hi[d] + 1) + 0.5) * prob.dx[d] for d in range(3)]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    rho = np.where(X < 0.4, 1.0, 0.125) * (1.0 + 0.05 * np.sin(2 * np.pi * (X + 2 * Y + 3 * Z)))
    p = np.where(X < 0.4, 1.0, 0.1) * (1.0 + 0.05 * np.cos(2 * np.pi * (2 * X - Y + Z)))
    u = 0.1 * np.sin(2 * np.pi * (Y + Z))
    v = 0.1 * np.cos(2 * np.pi * (X - Z))
    w = 0.1 * np.sin(2 * np.pi * (X + Y))
    S = np.zeros((10,) + rho.shape)
    S[0], S[1], S[2], S[3] = rho, rho * u, rho * v, rho * w
    S[5] = p / 0.4
    S[4] = S[5] + 0.5 * rho * (u * u + v * v + w * w)
    S[7] = rho
    cov = geom.covered_mask(g6lo, g6hi)
    S[:, cov] = 0.0
    S = np.ascontiguousarray(S)
    hp = hydro_params(prob)
    op = _oracle_params(oracle, prob)
    op.eb = 1
    dt = 0.2 * prob.dx[0]
    L3 = C.c_int * 3
    # oracle
    q = np.zeros((11,) + rho.shape)
    qa = np.zeros((5,) + rho.shape)
    err = C.c_int(0)
    oracle.lib().ora_ctoprim_box(L3(*g6lo), L3(*g6hi), C.byref(oracle.fab(S, g6lo)),
                                 C.byref(oracle.fab(q, g6lo)), C.byref(oracle.fab(qa, g6lo)),
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
    bg = geom.box_geom(b)
    g7 = tuple(x - 7 for x in lo)
    G = oracle.OraEbGeom()
    flags = np.ascontiguousarray(geom.flags, dtype=np.int32)
    G.flags = flags.ctypes.data_as(C.POINTER(C.c_int32))
    for d in range(3):
        G.flo[d], G.fhi[d] = int(geom.lo[d]), int(geom.hi[d])
    G.fjs = flags.shape[2]
    G.fks = flags.shape[2] * flags.shape[1]
    keep = {k: np.ascontiguousarray(v[None] if v.ndim == 3 else v) for k, v in bg.items()}
    G.vfrac = oracle.fab(keep["vfrac"], g7)
    for d, nm in enumerate("xyz"):
        G.ap[d] = oracle.fab(keep["ap" + nm], g7)
        G.fc[d] = oracle.fab(keep["fc" + nm], g7)
    srcq = np.zeros_like(q)
    oracle.lib().ora_hydro_umdrv_eb(L3(*lo), L3(*hi), C.byref(oracle.fab(S, g6lo)),
                                    C.byref(oracle.fab(dsdt, lo)), C.byref(oracle.fab(q, g6lo)),
                                    C.byref(oracle.fab(qa, g6lo)), C.byref(oracle.fab(srcq, g6lo)),
                                    fl, dt, C.byref(op), C.byref(G))
    # gpu
    dev = torch.device("cuda:0")
    tS = torch.from_numpy(S).to(dev)
    tq = torch.zeros(q.shape, dtype=torch.float64, device=dev)
    tqa = torch.zeros(qa.shape, dtype=torch.float64, device=dev)
    hydro_ctoprim(g6lo, g6hi, tS, g6lo, tq, tqa, hp)
    tds = torch.zeros(dsdt.shape, dtype=torch.float64, device=dev)
    tfl = [torch.zeros(f.shape, dtype=torch.float64, device=dev) for f in flx]
    tg = {k: torch.from_numpy(v if v.ndim == 4 else v).to(dev) for k, v in bg.items()}
    tg["flags"] = torch.from_numpy(flags).to(dev)
    tg["flo"] = tuple(int(v) for v in geom.lo)
    hydro_umdrv_eb(lo, hi, tS, tds, tq, tqa, tfl, tg, dt, hp, prob.is_periodic)
    torch.cuda.synchronize()
    assert np.array_equal(tq.cpu().numpy(), q)
    for d in range(3):
        assert np.array_equal(tfl[d].cpu().numpy(), flx[d]), d
    assert np.array_equal(tds.cpu().numpy(), dsdt)
    assert np.abs(dsdt).max() > 0.0


def test_eb_level_without_cut_cells_matches_plain_level():
    """A level of the EB build whose geometry has no cut cell (the sphere
    lies outside the domain) steps bit for bit like the plain level: the
    covered-cell guards and the 6-cell ghost region leave smooth data
    alone: regular boxes match the non-EB path."""
    from camr_amd import eb
    prob = workloads.smooth_periodic(n=16, ppm_type=0, cfl=0.4, max_grid=8)
    geom = eb.Geometry(prob, eb.Sphere((10.0, 10.0, 10.0), 0.1))
    A = Level(prob, engine=_lib.ENGINE_STAGED)
    A.set_eb(geom)
    B = Level(prob)
    for b in range(len(prob.boxes)):
        U = workloads.smooth_init_varp(prob, b)
        A.set_box(b, U)
        B.set_box(b, U)
    A.init_post()
    B.init_post()
    assert np.array_equal(np.asarray(A.run(3)), np.asarray(B.run(3)))
    for b in range(len(prob.boxes)):
        assert np.array_equal(A.get_box(b), B.get_box(b))


@pytest.mark.parametrize("n,max_grid", [(32, 8), (24, 8)])
def test_eb_merged_regular_boxes_match_box_by_box(n, max_grid):
    """The regular boxes of an EB level run as virtual boxes (blocks of equal
    boxes that tile a box, their Sborder filled as one row of expand_state)
    and box by box: the same dt sequence and the same state bit for bit, on
    Sod-past-a-sphere grids of 8^3 boxes (64x32x32; 48x24x24, whose boxes
    at the high faces are of other shapes)."""
    print(f"EB merge: n={n} max_grid={max_grid}, merged vs box by box", flush=True)
    prob = workloads.sod_sphere(n=n, max_grid=max_grid, interface=0.4)
    geom = workloads.sod_sphere_geometry(prob)
    A = Level(prob, engine=_lib.ENGINE_STAGED)
    A.set_eb(geom)
    B = Level(prob, engine=_lib.ENGINE_STAGED)
    B.set_eb(geom, merge=False)
    nv, nm = A.merge_info()
    assert nv > 0 and nm >= 2 * nv
    assert B.merge_info() == (0, 0)
    for b in range(len(prob.boxes)):
        U = workloads.sod_sphere_init(prob, b, geom)
        A.set_box(b, U)
        B.set_box(b, U)
    A.init_post()
    B.init_post()
    assert np.array_equal(np.asarray(A.run(5)), np.asarray(B.run(5)))
    for b in range(len(prob.boxes)):
        assert np.array_equal(A.get_box(b), B.get_box(b)), b
    assert np.array_equal(A.eb_force(), B.eb_force())
