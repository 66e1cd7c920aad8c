"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg only; the product never imports this module.  The oracle is a plain-C
restatement of CAMR's single-level Godunov advance (see camr_oracle.c for the
reference file:line each function follows).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CAMR_ORACLE_LIB selects another build of the oracle (the mutation tests
# load builds that carry a planted slip); the default is the committed recipe's.
_LIB = os.environ.get("CAMR_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")

NVAR, QVAR, NQAUX, NGDNV = 10, 11, 5, 6


class OraFab(C.Structure):
    _fields_ = [("p", C.POINTER(C.c_double)), ("lo", C.c_int * 3),
                ("hi", C.c_int * 3), ("js", C.c_int64), ("ks", C.c_int64),
                ("ns", C.c_int64), ("nc", C.c_int)]


class OraParams(C.Structure):
    _fields_ = [
        ("ppm_type", C.c_int), ("plm_iorder", C.c_int),
        ("use_pslope", C.c_int), ("use_flattening", C.c_int),
        ("transverse_reset_density", C.c_int),
        ("allow_negative_energy", C.c_int), ("allow_small_energy", C.c_int),
        ("dual_energy_update_E_from_e", C.c_int), ("add_grav_src", C.c_int),
        ("small", C.c_double), ("small_dens", C.c_double),
        ("small_pres", C.c_double), ("small_temp", C.c_double),
        ("smallu", C.c_double), ("small_num", C.c_double),
        ("difmag", C.c_double), ("dual_energy_eta1", C.c_double),
        ("dual_energy_eta2", C.c_double), ("const_grav", C.c_double),
        ("cfl", C.c_double), ("change_max", C.c_double),
        ("init_shrink", C.c_double), ("max_dt", C.c_double),
        ("fixed_dt", C.c_double), ("stop_time", C.c_double),
        ("bclo", C.c_int * 3), ("bchi", C.c_int * 3),
        ("domlo", C.c_int * 3), ("domhi", C.c_int * 3),
        ("is_periodic", C.c_int * 3), ("dx", C.c_double * 3),
        ("prob_lo", C.c_double * 3), ("tile", C.c_int * 3),
        ("eb", C.c_int), ("do_mol", C.c_int),
    ]


class OraEbGeom(C.Structure):
    _fields_ = [("flags", C.POINTER(C.c_int32)), ("flo", C.c_int * 3), ("fhi", C.c_int * 3),
                ("fjs", C.c_int64), ("fks", C.c_int64), ("vfrac", OraFab),
                ("ap", OraFab * 3), ("fc", OraFab * 3)]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"],
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        dp = C.POINTER(C.c_double)
        L.ora_default_params.argtypes = [C.POINTER(OraParams)]
        L.ora_riemannus.argtypes = [C.c_double] * 20 + [dp]
        L.ora_ppm_reconstruct.argtypes = [dp, C.c_double, dp, dp]
        L.ora_ppm_int_profile.argtypes = [C.c_double] * 6 + [dp, dp]
        L.ora_plm_slope.argtypes = [C.POINTER(OraFab)] + [C.c_int] * 5 + [C.c_double, C.c_int]
        L.ora_plm_slope.restype = C.c_double
        L.ora_flatten.argtypes = [C.POINTER(OraFab)] + [C.c_int] * 4
        L.ora_flatten.restype = C.c_double
        ip = C.POINTER(C.c_int)
        L.ora_ctoprim_box.argtypes = [ip, ip, C.POINTER(OraFab), C.POINTER(OraFab),
                                      C.POINTER(OraFab), C.POINTER(OraParams), ip]
        L.ora_srctoprim_box.argtypes = [ip, ip] + [C.POINTER(OraFab)] * 4
        L.ora_hydro_umdrv.argtypes = [ip, ip] + [C.POINTER(OraFab)] * 6 + [C.c_double, C.POINTER(OraParams)]
        L.ora_estdt_box.argtypes = [ip, ip, C.POINTER(OraFab), dp]
        L.ora_estdt_box.restype = C.c_double
        L.ora_level_create.argtypes = [C.POINTER(OraParams), C.c_int, ip]
        L.ora_level_create.restype = C.c_void_p
        L.ora_level_destroy.argtypes = [C.c_void_p]
        L.ora_level_set_exterior.argtypes = [C.c_void_p, dp]
        L.ora_level_set_box.argtypes = [C.c_void_p, C.c_int, dp]
        L.ora_level_get_box.argtypes = [C.c_void_p, C.c_int, dp]
        L.ora_level_init_post.argtypes = [C.c_void_p]
        L.ora_level_initial_dt.argtypes = [C.c_void_p]
        L.ora_level_initial_dt.restype = C.c_double
        L.ora_level_new_dt.argtypes = [C.c_void_p, C.c_double]
        L.ora_level_new_dt.restype = C.c_double
        L.ora_level_advance.argtypes = [C.c_void_p, C.c_double]
        L.ora_level_advance.restype = C.c_int
        L.ora_level_run.argtypes = [C.c_void_p, C.c_int, dp]
        L.ora_level_run.restype = C.c_int
        L.ora_level_time.argtypes = [C.c_void_p]
        L.ora_level_time.restype = C.c_double
        L.ora_level_sums.argtypes = [C.c_void_p, dp]
        ip32 = C.POINTER(C.c_int32)
        L.ora_level_set_eb.argtypes = [C.c_void_p, ip32, ip, ip]
        L.ora_level_set_eb.restype = C.c_int
        L.ora_hydro_umdrv_eb.argtypes = [ip, ip] + [C.POINTER(OraFab)] * 6 + [
            C.c_double, C.POINTER(OraParams), C.POINTER(OraEbGeom)]
        L.ora_level_box_kind.argtypes = [C.c_void_p, C.c_int]
        L.ora_level_box_kind.restype = C.c_int
        L.ora_level_set_box_eb.argtypes = [C.c_void_p, C.c_int] + [dp] * 7
        L.ora_level_exterior_from_coarse.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        L.ora_level_exterior_from_coarse.restype = C.c_int
        L.ora_level_get_exterior.argtypes = [C.c_void_p, dp]
        L.ora_fluxreg_create.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        L.ora_fluxreg_create.restype = C.c_void_p
        L.ora_fluxreg_destroy.argtypes = [C.c_void_p]
        L.ora_level_average_down.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        L.ora_level_average_down.restype = C.c_int
        L.ora_amr_step.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]
        L.ora_amr_step.restype = C.c_int
        L.ora_amr_advance.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int,
                                      C.c_double]
        L.ora_amr_advance.restype = C.c_int
        L.ora_level_eb_force.argtypes = [C.c_void_p, dp]
        L.ora_level_eb_force.restype = C.c_int
        L.ora_level_exterior_from_coarse_time.argtypes = [C.c_void_p, C.c_void_p, C.c_int] + [C.c_double] * 3
        L.ora_level_exterior_from_coarse_time.restype = C.c_int
        L.ora_level_set_old_box.argtypes = [C.c_void_p, C.c_int, dp]
        F = C.POINTER(OraFab)
        L.ora_transdo_point.argtypes = [C.c_int] * 5 + [F] * 7 + [C.c_double, C.c_int, C.c_double]
        L.ora_transdd_point.argtypes = [C.c_int] * 4 + [F] * 10 + [C.c_double] * 3 + [C.c_int, C.c_double]
        L.ora_trace_ppm_point.argtypes = [C.c_int] * 4 + [F] * 4 + [C.c_double, dp, C.c_int,
                                                                    C.c_double, C.c_double]
        L.ora_artif_visc_point.argtypes = [C.c_int] * 3 + [F] * 3 + [C.c_double, C.c_double] + [C.c_int] * 5
        L.ora_divu_node_point.argtypes = [C.c_int] * 3 + [F, F, C.POINTER(OraParams)]
        L.ora_enforce_min_density_box.argtypes = [F, ip, ip, C.POINTER(OraParams)]
        L.ora_enforce_min_density_box.restype = C.c_int
        L.ora_set_threads.argtypes = [C.c_int]
        L.ora_get_threads.restype = C.c_int
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def fab(arr: np.ndarray, lo) -> OraFab:
    """View a (nc, nz, ny, nx) float64 C-contiguous array as an ora_fab."""
    assert arr.dtype == np.float64 and arr.flags.c_contiguous and arr.ndim == 4
    nc, nz, ny, nx = arr.shape
    f = OraFab()
    f.p = _dp(arr)
    for d in range(3):
        f.lo[d] = int(lo[d])
    f.hi[0], f.hi[1], f.hi[2] = lo[0] + nx - 1, lo[1] + ny - 1, lo[2] + nz - 1
    f.js, f.ks, f.ns, f.nc = nx, nx * ny, nx * ny * nz, nc
    return f


def params(**kw) -> OraParams:
    p = OraParams()
    lib().ora_default_params(C.byref(p))
    for k, v in kw.items():
        if isinstance(getattr(p, k), C.Array):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(p, k, v)
    return p


def riemannus(rl, ul, vl, v2l, pl, rel, rr, ur, vr, v2r, pr, rer, cav,
              gamcl, gamcr, small, smallc, small_dens, small_pres, smallu):
    out = np.zeros(12)
    lib().ora_riemannus(rl, ul, vl, v2l, pl, rel, rr, ur, vr, v2r, pr, rer,
                        cav, gamcl, gamcr, small, smallc, small_dens,
                        small_pres, smallu, _dp(out))
    return out


def ppm_reconstruct(s, flat):
    s = np.ascontiguousarray(s, dtype=np.float64)
    sm, sp = C.c_double(), C.c_double()
    lib().ora_ppm_reconstruct(_dp(s), flat, C.byref(sm), C.byref(sp))
    return sm.value, sp.value


def ppm_int_profile(sm, sp, sc, u, c, dtdx):
    Ip, Im = np.zeros(3), np.zeros(3)
    lib().ora_ppm_int_profile(sm, sp, sc, u, c, dtdx, _dp(Ip), _dp(Im))
    return Ip, Im


class Level:
    """Single-level CAMR_advance restatement over a list of boxes."""

    def __init__(self, p: OraParams, boxes):
        self.p = p
        self.boxes = [tuple(map(int, b)) for b in boxes]
        lohi = np.array([x for b in self.boxes for x in b], dtype=np.int32)
        self._h = lib().ora_level_create(C.byref(p), len(self.boxes),
                                         lohi.ctypes.data_as(C.POINTER(C.c_int)))

    def close(self):
        if self._h:
            lib().ora_level_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, b):
        lo, hi = self.boxes[b][:3], self.boxes[b][3:]
        return (NVAR, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)

    def set_box(self, b, data):
        a = np.ascontiguousarray(data, dtype=np.float64)
        assert a.shape == self.shape(b)
        lib().ora_level_set_box(self._h, b, _dp(a))

    def get_box(self, b):
        a = np.empty(self.shape(b))
        lib().ora_level_get_box(self._h, b, _dp(a))
        return a

    def set_eb(self, geom):
        """EB build of the level with the cut-cell data of `geom`
        (``camr_amd.eb.Geometry``); call before set_box."""
        flags = np.ascontiguousarray(geom.flags, dtype=np.int32)
        flo = np.array(geom.lo, dtype=np.int32)
        fhi = np.array(geom.hi, dtype=np.int32)
        ipt = C.POINTER(C.c_int)
        lib().ora_level_set_eb(self._h, flags.ctypes.data_as(C.POINTER(C.c_int32)),
                               flo.ctypes.data_as(ipt), fhi.ctypes.data_as(ipt))
        self.kinds = [lib().ora_level_box_kind(self._h, b) for b in range(len(self.boxes))]
        for b, k in enumerate(self.kinds):
            if k == 1:
                g = geom.box_geom(b)
                arrs = [np.ascontiguousarray(g[n], dtype=np.float64)
                        for n in ("vfrac", "apx", "apy", "apz", "fcx", "fcy", "fcz")]
                lib().ora_level_set_box_eb(self._h, b, *[_dp(a) for a in arrs])
        self._geom_keep = geom

    def exterior_from_coarse(self, coarse, ratio):
        """Exterior data interpolated from level `coarse` (FillPatch's
        coarse-fine interpolation)."""
        assert lib().ora_level_exterior_from_coarse(self._h, coarse._h, int(ratio)) == 0

    def eb_force(self):
        f = np.zeros(3)
        assert lib().ora_level_eb_force(self._h, _dp(f)) == 0
        return f

    def exterior_from_coarse_time(self, coarse, ratio, t_old, t_new, t):
        assert lib().ora_level_exterior_from_coarse_time(self._h, coarse._h, ratio, t_old, t_new, t) == 0

    def set_old_box(self, b, data):
        data = np.ascontiguousarray(data, dtype=np.float64)
        assert data.shape == self.shape(b)
        lib().ora_level_set_old_box(self._h, b, _dp(data))

    def get_exterior(self):
        n = [self.p.domhi[d] - self.p.domlo[d] + 1 for d in range(3)]
        a = np.empty((NVAR, n[2], n[1], n[0]))
        lib().ora_level_get_exterior(self._h, _dp(a))
        return a

    def set_exterior(self, data):
        n = [self.p.domhi[d] - self.p.domlo[d] + 1 for d in range(3)]
        a = np.ascontiguousarray(data, dtype=np.float64)
        assert a.shape == (NVAR, n[2], n[1], n[0])
        lib().ora_level_set_exterior(self._h, _dp(a))

    def init_post(self):
        lib().ora_level_init_post(self._h)

    def initial_dt(self):
        return lib().ora_level_initial_dt(self._h)

    def new_dt(self, dt_old):
        return lib().ora_level_new_dt(self._h, dt_old)

    def advance(self, dt):
        return lib().ora_level_advance(self._h, dt)

    def run(self, nsteps):
        dts = np.zeros(max(nsteps, 1))
        err = lib().ora_level_run(self._h, nsteps, _dp(dts))
        return err, dts[:nsteps]

    def sums(self):
        s = np.zeros(NVAR)
        lib().ora_level_sums(self._h, _dp(s))
        return s

    @property
    def time(self):
        return lib().ora_level_time(self._h)


class FluxRegister:
    """Flux register between coarse level C and fine level F (YAFluxRegister
    restated); attaches itself to both levels."""

    def __init__(self, C: Level, F: Level, ratio: int):
        self.C, self.F, self.ratio = C, F, ratio
        self._h = lib().ora_fluxreg_create(C._h, F._h, ratio)
        if not self._h:
            raise ValueError("flux register: unsupported level pair")

    def close(self):
        if self._h:
            lib().ora_fluxreg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def average_down(C: Level, F: Level, ratio: int):
    assert lib().ora_level_average_down(C._h, F._h, ratio) == 0


def amr_step(C: Level, F: Level, R: FluxRegister, dt: float) -> int:
    """One subcycled two-level step of coarse size dt (fine steps dt/ratio)."""
    return lib().ora_amr_step(C._h, F._h, R._h, dt)


def amr_advance(levels, regs, dt0: float) -> int:
    """One coarse step of a hierarchy (levels coarse first, regs[l] between
    levels l and l+1) with subcycling at every level."""
    n = len(levels)
    hl = (C.c_void_p * n)(*[lv._h for lv in levels])
    hr = (C.c_void_p * max(n - 1, 1))(*([r._h for r in regs] or [None]))
    return lib().ora_amr_advance(hl, hr, n, dt0)


def set_threads(n: int):
    lib().ora_set_threads(int(n))
