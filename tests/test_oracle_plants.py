"""Mutation tests of the oracle's pins.

The oracle (oracle/camr_oracle.c) carries planted mistakes behind
ORA_PLANT; the default build has none.  Each case below builds the oracle
with one slip and runs the pins that must catch it in a fresh process
against that build (CAMR_ORACLE_LIB); the case passes when they fail.
Every slip is one a restatement of the Godunov composite can plausibly
make, and which the whole-step invariants (free stream, conservation) let
through.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# slip -> (what it is, the pins that catch it)
PLANTS = {
    1: ("transdo pressure correction with the wrong sign (Godunov_utils_3D.H:169,225)",
        "transdo_is_a_conservative"),
    2: ("transdo corrections with hdt = dt/2dx instead of cdtdx = dt/3dx "
        "(Godunov_3D.cpp:50-56,287)", "exact_swept_volume"),
    3: ("transdd of direction d fed the flux of t0 corrected by d instead of by t1 "
        "(a swapped transverse pair, Godunov_3D.cpp:393-439)", "exact_swept_volume"),
    4: ("PPM plus-state normal velocity with the acoustic projections swapped "
        "(PPM.cpp:291)", "ppm_projection"),
    5: ("artificial viscosity on expansion instead of compression "
        "(hydro_artif_visc.H:33-36)", "artificial_viscosity"),
    6: ("density floor ties resolved to the last neighbour (CAMR.cpp:1006)",
        "ties_go_to_the_first"),
    7: ("transdo pressure law with the Godunov velocity of the state's own direction "
        "(Godunov_utils_3D.H:49)", "transdo_is_a_conservative"),
    8: ("transdd pressure correction of the second transverse direction dropped "
        "(Godunov_utils_3D.H:384)", "transdd_combines"),
}


@pytest.fixture(scope="module")
def plant_dir(tmp_path_factory):
    out = tmp_path_factory.mktemp("oracle_plants")
    procs = [subprocess.Popen(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "plant",
                               f"N={n}", f"OUT={out}"]) for n in PLANTS]
    assert all(p.wait() == 0 for p in procs)
    return out


@pytest.mark.parametrize("n", sorted(PLANTS))
def test_planted_slip_is_caught(plant_dir, n):
    what, pins = PLANTS[n]
    env = dict(os.environ, CAMR_ORACLE_LIB=str(plant_dir / f"liboracle_plant{n}.so"))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_godunov.py"), "-k", pins],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 1 and " failed" in r.stdout, (what, r.stdout[-2000:])
