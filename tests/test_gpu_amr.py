"""Two-level subcycled steps with flux registers on the GPU against the
oracle, bit for bit: CrseAdd on the coarse level's per-box fluxes, FineAdd
on the fine level's, the fine ghost data interpolated from the coarse
level in time and space, reflux, average_down and the coarse
post_timestep clean (camr_amr_step), on Godunov PPM / PLM and the method
of lines, periodic and walled domains, ratios 2 and 4."""
import numpy as np
import pytest

from camr_amd import _lib
from camr_amd.hydro import FluxRegister, Level, Problem, amr_step
from tests.test_oracle_amr import FINE2, _boxes, _field, composite_sums

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FINE4 = [(16, 16, 16, 31, 47, 47), (32, 16, 16, 47, 47, 47)]
# fine boxes on the domain faces (x low and z high): the ghost shell wraps
# into periodic images or leaves a walled domain, and the registers border
# the faces
FACE = {2: [(0, 8, 8, 15, 23, 31)], 4: [(0, 16, 16, 31, 47, 63)]}


def _pair(O, n, boxes, periodic, kw, graph=False, engine=_lib.ENGINE_STAGED):
    bc = 0 if periodic else 4
    per = (1, 1, 1) if periodic else (0, 0, 0)
    params = dict(cfl=0.3, **kw)
    op = O.params(domlo=(0, 0, 0), domhi=(n - 1,) * 3, dx=(1.0 / n,) * 3, bclo=(bc,) * 3,
                  bchi=(bc,) * 3, is_periodic=per, **params)
    prob = Problem(domlo=(0, 0, 0), domhi=(n - 1,) * 3, dx=(1.0 / n,) * 3, bclo=(bc,) * 3,
                   bchi=(bc,) * 3, is_periodic=per, boxes=boxes, params=params)
    ol = O.Level(op, boxes)
    gl = Level(prob, engine=engine, graph=graph)
    for b, bb in enumerate(boxes):
        U = _field(n, bb[:3], bb[3:])
        ol.set_box(b, U)
        gl.set_box(b, U)
    return ol, gl


FUSED = _lib.ENGINE_FUSED


@pytest.mark.parametrize("ratio,periodic,kw,graph,engine", [
    (2, True, dict(ppm_type=1), False, _lib.ENGINE_STAGED),
    (2, False, dict(ppm_type=0), False, _lib.ENGINE_STAGED),
    (2, True, dict(do_mol=1), False, _lib.ENGINE_STAGED),
    (2, False, dict(do_mol=1), False, _lib.ENGINE_STAGED),
    (4, True, dict(ppm_type=1), False, _lib.ENGINE_STAGED),
    (4, False, dict(ppm_type=1), False, _lib.ENGINE_STAGED),
    (2, True, dict(ppm_type=1), True, _lib.ENGINE_STAGED),
    (2, True, dict(do_mol=1), True, _lib.ENGINE_STAGED),
    (2, True, dict(ppm_type=1), False, FUSED),
    (2, False, dict(ppm_type=0), False, FUSED),
    (4, True, dict(ppm_type=1), False, FUSED),
    (4, False, dict(ppm_type=1), False, FUSED),
    (2, True, dict(ppm_type=1, face=1), False, _lib.ENGINE_STAGED),
    (2, False, dict(ppm_type=0, face=1), False, _lib.ENGINE_STAGED),
    (4, True, dict(ppm_type=1, face=1), False, _lib.ENGINE_STAGED),
    (4, False, dict(ppm_type=1, face=1), False, _lib.ENGINE_STAGED),
    (2, True, dict(ppm_type=1, face=1), False, FUSED),
    (2, False, dict(ppm_type=0, face=1), False, FUSED),
    (4, True, dict(ppm_type=1, face=1), False, FUSED),
    (4, False, dict(ppm_type=1, face=1), False, FUSED),
])
def test_amr_step_matches_oracle(oracle, ratio, periodic, kw, graph, engine):
    """graph: every level step captured (per buffer parity and deferred
    clean) and replayed, the registers' adds inside the graphs.  FUSED: the
    fused engine's canonical flux arrays feed the registers.  face: the fine
    boxes touch domain faces."""
    print(f"two-level step ratio={ratio} periodic={periodic} {kw} graph={graph} "
          f"engine={engine}", flush=True)
    nc = 16
    kw = dict(kw)
    if kw.pop("face", 0):
        fine = FACE[ratio]
    else:
        fine = FINE2 if ratio == 2 else FINE4
    oc, gc = _pair(oracle, nc, _boxes(nc, 8), periodic, kw, graph, engine)
    of, gf = _pair(oracle, nc * ratio, fine, periodic, kw, graph, engine)
    oR = oracle.FluxRegister(oc, of, ratio)
    gR = FluxRegister(gc, gf, ratio)
    oracle.average_down(oc, of, ratio)
    gc.average_down(gf, ratio)
    for L in (oc, of, gc, gf):
        L.init_post()
    s0 = composite_sums(gc, gf, ratio, nc * ratio)
    dt = None
    for s in range(4):
        if s == 0:
            odt = min(oc.initial_dt(), ratio * of.initial_dt())
            gdt = min(gc.initial_dt(), ratio * gf.initial_dt())
        else:
            odt = min(oc.new_dt(dt), ratio * of.new_dt(dt / ratio))
            gdt = min(gc.new_dt(dt), ratio * gf.new_dt(dt / ratio))
        assert odt == gdt, (s, odt, gdt)
        dt = odt
        assert oracle.amr_step(oc, of, oR, dt) == 0
        amr_step(gc, gf, gR, dt)
        for b in range(len(oc.boxes)):
            assert np.array_equal(gc.get_box(b), oc.get_box(b)), ("coarse", s, b)
        for b in range(len(of.boxes)):
            assert np.array_equal(gf.get_box(b), of.get_box(b)), ("fine", s, b)
    if periodic:
        s1 = composite_sums(gc, gf, ratio, nc * ratio)
        for c in (0, 1, 2, 3, 4):
            assert abs(s1[c] - s0[c]) <= 1e-12 * max(abs(s0[c]), 1.0)
    gR.close()
    oR.close()


def test_register_refuses_unsupported_layouts():
    prob_c = Problem(domlo=(0, 0, 0), domhi=(15,) * 3, dx=(1 / 16,) * 3, bclo=(0,) * 3,
                     bchi=(0,) * 3, is_periodic=(1, 1, 1), boxes=_boxes(16, 8),
                     params=dict(cfl=0.3))

    def fine(boxes):
        return Problem(domlo=(0, 0, 0), domhi=(31,) * 3, dx=(1 / 32,) * 3, bclo=(0,) * 3,
                       bchi=(0,) * 3, is_periodic=(1, 1, 1), boxes=boxes, params=dict(cfl=0.3))

    gc = Level(prob_c, engine=_lib.ENGINE_STAGED)
    # a coarse cell bordering two fine boxes
    gf = Level(fine([(8, 8, 8, 23, 23, 23), (24, 8, 8, 31, 15, 23)]), engine=_lib.ENGINE_STAGED)
    with pytest.raises(_lib.CamrError, match="borders two fine boxes"):
        FluxRegister(gc, gf, 2)
    # fine boxes not aligned to the ratio
    gf = Level(fine([(9, 8, 8, 23, 23, 23)]), engine=_lib.ENGINE_STAGED)
    with pytest.raises(_lib.CamrError, match="aligned"):
        FluxRegister(gc, gf, 2)


@pytest.mark.parametrize("engine", [_lib.ENGINE_STAGED, FUSED])
def test_three_level_advance_matches_oracle(oracle, engine):
    """Three levels (16^3, a 2x block at ratio 2, a nested block at ratio 2
    inside it), two coarse steps of recursive subcycling: every level bit
    for bit, the composite sums conserved."""
    from camr_amd.hydro import amr_advance
    from tests.test_oracle_amr import FINE3, composite_sums3
    kw = dict(ppm_type=1)
    pairs = [_pair(oracle, 16, _boxes(16, 8), True, kw, engine=engine),
             _pair(oracle, 32, FINE2, True, kw, engine=engine),
             _pair(oracle, 64, FINE3, True, kw, engine=engine)]
    ol = [p[0] for p in pairs]
    gl = [p[1] for p in pairs]
    oR = [oracle.FluxRegister(ol[0], ol[1], 2), oracle.FluxRegister(ol[1], ol[2], 2)]
    gR = [FluxRegister(gl[0], gl[1], 2), FluxRegister(gl[1], gl[2], 2)]
    for Ls, avg in ((ol, lambda c, f: oracle.average_down(c, f, 2)),
                    (gl, lambda c, f: c.average_down(f, 2))):
        avg(Ls[1], Ls[2])
        avg(Ls[0], Ls[1])
        for L in Ls:
            L.init_post()
    s0 = composite_sums3(gl)
    dt = None
    for s in range(2):
        oests = [L.initial_dt() if s == 0 else L.new_dt(dt / 2 ** l) for l, L in enumerate(ol)]
        gests = [L.initial_dt() if s == 0 else L.new_dt(dt / 2 ** l) for l, L in enumerate(gl)]
        assert oests == gests
        dt = min(e * 2 ** l for l, e in enumerate(oests))
        assert oracle.amr_advance(ol, oR, dt) == 0
        amr_advance(gl, gR, dt)
        for l in range(3):
            for b in range(len(ol[l].boxes)):
                assert np.array_equal(gl[l].get_box(b), ol[l].get_box(b)), (s, l, b)
    s1 = composite_sums3(gl)
    for c in (0, 1, 2, 3, 4):
        assert abs(s1[c] - s0[c]) <= 1e-12 * max(abs(s0[c]), 1.0)
    for r in gR + oR:
        r.close()


@pytest.mark.parametrize("engine", [_lib.ENGINE_STAGED, FUSED])
def test_density_floor_inside_a_two_level_step(oracle, engine):
    """Cells below small_dens on both levels: enforce_min_density fires in
    the composite step (on the fused engine the step is redone exactly
    before the register adds, so they are counted once)."""
    kw = dict(ppm_type=0)
    oc, gc = _pair(oracle, 16, _boxes(16, 8), True, kw, engine=engine)
    of, gf = _pair(oracle, 32, FINE2, True, kw, engine=engine)

    def plant(L, b, cells):
        U = L.get_box(b)
        bb = L.boxes[b]
        for (i, j, k), rho in cells:
            ii, jj, kk = i - bb[0], j - bb[1], k - bb[2]
            f = rho / U[0, kk, jj, ii]
            U[[0, 1, 2, 3, 5, 7, 8, 9], kk, jj, ii] *= f
            U[4, kk, jj, ii] = U[5, kk, jj, ii] + 0.5 * (U[1:4, kk, jj, ii] ** 2).sum() / U[0, kk, jj, ii]
        return U

    for (o, g) in ((oc, gc), (of, gf)):
        b = 0
        cells = [((o.boxes[b][0] + 1, o.boxes[b][1] + 1, o.boxes[b][2] + 1), 5e-8),
                 ((o.boxes[b][0] + 2, o.boxes[b][1] + 1, o.boxes[b][2] + 1), 2e-8)]
        U = plant(o, b, cells)
        o.set_box(b, U)
        g.set_box(b, U)
    oR = oracle.FluxRegister(oc, of, 2)
    gR = FluxRegister(gc, gf, 2)
    oracle.average_down(oc, of, 2)
    gc.average_down(gf, 2)
    for L in (oc, of, gc, gf):
        L.init_post()
    dt = None
    for s in range(2):
        odt = min(oc.initial_dt(), 2 * of.initial_dt()) if s == 0 else \
            min(oc.new_dt(dt), 2 * of.new_dt(dt / 2))
        gdt = min(gc.initial_dt(), 2 * gf.initial_dt()) if s == 0 else \
            min(gc.new_dt(dt), 2 * gf.new_dt(dt / 2))
        assert odt == gdt
        dt = odt
        assert oracle.amr_step(oc, of, oR, dt) == 0
        amr_step(gc, gf, gR, dt)
        for b in range(len(oc.boxes)):
            assert np.array_equal(gc.get_box(b), oc.get_box(b)), ("coarse", s, b)
        for b in range(len(of.boxes)):
            assert np.array_equal(gf.get_box(b), of.get_box(b)), ("fine", s, b)
    gR.close()
    oR.close()
