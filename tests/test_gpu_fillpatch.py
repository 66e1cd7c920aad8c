"""Coarse-to-fine ghost data on the GPU (camr_level_exterior_from_coarse)
against the oracle: a refined level whose ghost cells outside its boxes come
from the coarse level's state by FillPatch's conservative interpolation
steps bit for bit like the oracle's, on both engines."""
import numpy as np
import pytest

from camr_amd import _lib, workloads
from camr_amd.hydro import Level

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.mark.parametrize("engine", [_lib.ENGINE_FUSED, _lib.ENGINE_STAGED])
def test_refined_level_from_coarse_matches_oracle(oracle, engine):
    from test_gpu_parity import _oracle_params
    coarse = workloads.smooth_periodic(n=16, ppm_type=1, cfl=0.4, max_grid=8)
    fine = workloads.refined_level(n=32, radius=0.3, blocking_factor=8, max_size=16)
    print(f"fine boxes {len(fine.boxes)} engine {engine}", flush=True)
    oc = oracle.Level(_oracle_params(oracle, coarse), coarse.boxes)
    gc = Level(coarse)
    for b in range(len(coarse.boxes)):
        U = workloads.smooth_init(coarse, b)
        oc.set_box(b, U)
        gc.set_box(b, U)
    oc.init_post()
    gc.init_post()
    err, odt = oc.run(1)
    assert err == 0
    assert np.array_equal(np.asarray(gc.run(1)), odt)
    of = oracle.Level(_oracle_params(oracle, fine), fine.boxes)
    gf = Level(fine, engine=engine)
    for b in range(len(fine.boxes)):
        U = workloads.smooth_init(fine, b)
        of.set_box(b, U)
        gf.set_box(b, U)
    of.exterior_from_coarse(oc, 2)
    gf.exterior_from_coarse(gc, 2)
    of.init_post()
    gf.init_post()
    err, odts = of.run(2)
    assert err == 0
    assert np.array_equal(np.asarray(gf.run(2)), odts)
    for b in range(len(fine.boxes)):
        assert np.array_equal(gf.get_box(b), of.get_box(b)), b
