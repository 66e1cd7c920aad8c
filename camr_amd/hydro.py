"""Host-side mirror of CAMR's level advance over the gfx950 C-ABI.

``Level`` drives ``camr_level_*`` (the CAMR_advance / computeNewDt mirror in
camr_amd/csrc/camr_level.cpp); ``hydro_umdrv`` and friends call the per-box
drop-in entry points on torch device tensors.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

# CAMR parameter defaults (Source/Params/_cpp_parameters, CAMR_Constants.H)
DEFAULTS = dict(ppm_type=0, plm_iorder=4, use_pslope=0, use_flattening=1,
                transverse_reset_density=1, allow_negative_energy=0,
                allow_small_energy=1, dual_energy_update_E_from_e=1,
                small=1.e-8, small_dens=1.e-7, small_pres=1.e-200,
                small_temp=1.e2, smallu=1.e-12, small_num=1.e-8,
                difmag=0.1, dual_energy_eta1=1.0, dual_energy_eta2=1.0e-4,
                do_mol=0)
LEVEL_DEFAULTS = dict(add_grav_src=1, const_grav=0.0, cfl=0.8,
                      change_max=1.1, init_shrink=1.0, max_dt=1.e200,
                      fixed_dt=-1.0, stop_time=-1.0)


@dataclass
class Problem:
    """A single-level problem: domain, BCs, boxes and runtime parameters."""
    domlo: tuple
    domhi: tuple
    dx: tuple
    bclo: tuple
    bchi: tuple
    is_periodic: tuple
    boxes: list
    prob_lo: tuple = (0.0, 0.0, 0.0)
    params: dict = field(default_factory=dict)

    def value(self, k):
        if k in self.params:
            return self.params[k]
        return DEFAULTS.get(k, LEVEL_DEFAULTS.get(k))


def hydro_params(prob: Problem) -> L.HydroParams:
    p = L.HydroParams()
    L.lib().camr_default_params(C.byref(p))
    for k in DEFAULTS:
        setattr(p, k, prob.value(k))
    for d in range(3):
        p.bclo[d], p.bchi[d] = prob.bclo[d], prob.bchi[d]
        p.domlo[d], p.domhi[d] = prob.domlo[d], prob.domhi[d]
        p.dx[d] = prob.dx[d]
    return p


class Level:
    """One AMR level advanced on the GPU by the C-ABI level driver."""

    def __init__(self, prob: Problem, engine: int = L.ENGINE_AUTO, batched: bool = True,
                 graph: bool = False, worker_streams: int = 0):
        self.prob = prob
        d = L.LevelDesc()
        L.lib().camr_default_level_desc(C.byref(d))
        d.hp = hydro_params(prob)
        for dd in range(3):
            d.is_periodic[dd] = int(prob.is_periodic[dd])
        for k in LEVEL_DEFAULTS:
            setattr(d, k, prob.value(k))
        self._lohi = np.ascontiguousarray(
            np.array([x for b in prob.boxes for x in b], dtype=np.int32))
        d.nbox = len(prob.boxes)
        d.box_lohi = self._lohi.ctypes.data_as(C.POINTER(C.c_int32))
        d.engine = engine
        d.batched = int(bool(batched))
        d.graph = int(bool(graph))
        if worker_streams:
            d.worker_streams = int(worker_streams)
        self._desc = d
        h = C.c_void_p()
        L.check(L.lib().camr_level_create(C.byref(d), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            L.lib().camr_level_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, b):
        lo, hi = self.prob.boxes[b][:3], self.prob.boxes[b][3:]
        return (L.NVAR, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)

    def set_box(self, b, data):
        a = np.ascontiguousarray(data, dtype=np.float64)
        assert a.shape == self.shape(b)
        L.check(L.lib().camr_level_upload_box(
            self._h, b, a.ctypes.data_as(C.POINTER(C.c_double))))

    @property
    def boxes(self):
        return self.prob.boxes

    def get_box(self, b):
        a = np.empty(self.shape(b))
        L.check(L.lib().camr_level_download_box(
            self._h, b, a.ctypes.data_as(C.POINTER(C.c_double))))
        return a

    def exterior_from_coarse(self, coarse: "Level", ratio: int):
        """Ghost data outside the level's boxes from the coarser level
        `coarse` (FillPatch's conservative linear interpolation), computed on
        the device from its current state."""
        L.check(L.lib().camr_level_exterior_from_coarse(self._h, coarse._h, int(ratio)))

    def eb_force(self):
        """The force on the embedded body from the current state: the wall
        momentum flux of every cut cell times dx^2, summed (x, y, z)."""
        f = (C.c_double * 3)()
        L.check(L.lib().camr_level_eb_force(self._h, f))
        return np.array(list(f))

    def average_down(self, fine: "Level", ratio: int):
        """CAMR::avgDown: the coarse cells covered by `fine` become the mean
        of their children."""
        L.check(L.lib().camr_level_average_down(self._h, fine._h, int(ratio)))

    def set_eb(self, geom, merge: bool = True):
        """Embedded-boundary build of the level with the cut-cell data of
        `geom` (``camr_amd.eb.Geometry``).  The level must have been created
        with ``engine=ENGINE_STAGED``; call before the first step.  `merge`:
        equal regular boxes that tile a box in blocks of up to 4x4x4 run as
        that box (same results, fewer ghost cells)."""
        i32 = C.POINTER(C.c_int32)
        L.check(L.lib().camr_level_set_eb_merge(self._h, 1 if merge else 0))
        flags = np.ascontiguousarray(geom.flags, dtype=np.int32)
        flo = np.ascontiguousarray(geom.lo, dtype=np.int32)
        fhi = np.ascontiguousarray(geom.hi, dtype=np.int32)
        L.check(L.lib().camr_level_set_eb(self._h, flags.ctypes.data_as(i32),
                                          flo.ctypes.data_as(i32), fhi.ctypes.data_as(i32)))
        self.kinds = []
        for b in range(len(self.prob.boxes)):
            k = C.c_int32()
            L.check(L.lib().camr_level_box_kind(self._h, b, C.byref(k)))
            self.kinds.append(k.value)
            if k.value == 1:
                g = geom.box_geom(b)
                arrs = [np.ascontiguousarray(g[n], dtype=np.float64)
                        for n in ("vfrac", "apx", "apy", "apz", "fcx", "fcy", "fcz")]
                L.check(L.lib().camr_level_set_box_eb(
                    self._h, b, *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]))

    def merge_info(self):
        """(virtual boxes, level boxes they hold) of an EB level's batch of
        regular boxes."""
        nv, nm = C.c_int32(), C.c_int32()
        L.check(L.lib().camr_level_eb_merge_info(self._h, C.byref(nv), C.byref(nm)))
        return nv.value, nm.value

    def state_array(self, b) -> L.Array4:
        a = L.Array4()
        L.check(L.lib().camr_level_state_array(self._h, b, C.byref(a)))
        return a

    def set_stream(self, stream_ptr: int):
        L.check(L.lib().camr_level_set_stream(self._h, C.c_void_p(stream_ptr)))

    def init_post(self):
        L.check(L.lib().camr_level_init_post(self._h))

    def initial_dt(self):
        v = C.c_double()
        L.check(L.lib().camr_level_initial_dt(self._h, C.byref(v)))
        return v.value

    def new_dt(self, dt_old):
        v = C.c_double()
        L.check(L.lib().camr_level_new_dt(self._h, dt_old, C.byref(v)))
        return v.value

    def advance(self, dt):
        L.check(L.lib().camr_level_advance(self._h, dt))

    def sync(self):
        L.check(L.lib().camr_level_sync(self._h))

    def run(self, nsteps):
        dts = np.zeros(max(nsteps, 1))
        L.check(L.lib().camr_level_run(
            self._h, nsteps, dts.ctypes.data_as(C.POINTER(C.c_double))))
        return dts[:nsteps]

    @property
    def time(self):
        return L.lib().camr_level_time(self._h)

    # ------------------------------------------------------------ adjoint
    def adjoint_forward(self, dts, every=5):
        """From the current state S_0, len(dts) steps with these (frozen) dts,
        keeping device checkpoints every `every` steps; ends at S_T."""
        d = np.ascontiguousarray(dts, dtype=np.float64)
        L.check(L.lib().camr_level_adjoint_forward(
            self._h, len(d), d.ctypes.data_as(C.POINTER(C.c_double)), int(every)))

    def set_adjoint(self, b, data):
        """dL/dS_T of box b, (NVAR, nz, ny, nx)."""
        a = np.ascontiguousarray(data, dtype=np.float64)
        assert a.shape == self.shape(b)
        L.check(L.lib().camr_level_adj_upload_box(
            self._h, b, a.ctypes.data_as(C.POINTER(C.c_double))))

    def get_adjoint(self, b):
        a = np.empty(self.shape(b))
        L.check(L.lib().camr_level_adj_download_box(
            self._h, b, a.ctypes.data_as(C.POINTER(C.c_double))))
        return a

    def adjoint_backward(self):
        """dL/dS_T (set_adjoint) -> dL/dS_0 (get_adjoint); the level ends at S_0."""
        L.check(L.lib().camr_level_adjoint_backward(self._h))

    def seed_force_adjoint(self, w=(1.0, 0.0, 0.0)):
        """After adjoint_forward on an embedded-boundary level: dL/dS_T for
        L = w . F, F the force on the body at the final state (eb_force)."""
        a = np.ascontiguousarray(w, dtype=np.float64)
        L.check(L.lib().camr_level_adj_seed_eb_force(
            self._h, a.ctypes.data_as(C.POINTER(C.c_double))))

    def set_exterior(self, data):
        """Values for ghost cells no box owns: (NVAR, nz, ny, nx) over the domain."""
        n = [self.prob.domhi[d] - self.prob.domlo[d] + 1 for d in range(3)]
        a = np.ascontiguousarray(data, dtype=np.float64)
        assert a.shape == (L.NVAR, n[2], n[1], n[0])
        L.check(L.lib().camr_level_set_exterior(self._h, a.ctypes.data_as(C.POINTER(C.c_double))))

    @property
    def tuning(self):
        """(block sizes of the CAMR_NSTAGE stage ranges, box order)."""
        blk = (C.c_int32 * L.NSTAGE)()
        order = C.c_int32()
        L.check(L.lib().camr_level_get_tuning(self._h, blk, C.byref(order)))
        return list(blk), order.value

    @tuning.setter
    def tuning(self, t):
        blk, order = t
        arr = (C.c_int32 * L.NSTAGE)(*blk)
        L.check(L.lib().camr_level_set_tuning(self._h, arr, int(order)))

    @property
    def tiles(self):
        """Block shape (tx, ty, tz) of each stage range; (0, 0, 0) = blocks of
        consecutive points in x-fastest order."""
        t = (C.c_int32 * (3 * L.NSTAGE))()
        L.check(L.lib().camr_level_get_tiles(self._h, t))
        return [tuple(t[3 * s:3 * s + 3]) for s in range(L.NSTAGE)]

    @tiles.setter
    def tiles(self, tiles):
        flat = [int(x) for tt in tiles for x in tt]
        L.check(L.lib().camr_level_set_tiles(self._h, (C.c_int32 * (3 * L.NSTAGE))(*flat)))

    @property
    def xcd(self):
        """XCD chunk of each stage range (0: the hardware's block order)."""
        t = (C.c_int32 * L.NSTAGE)()
        L.check(L.lib().camr_level_get_xcd(self._h, t))
        return list(t)

    @xcd.setter
    def xcd(self, chunks):
        L.check(L.lib().camr_level_set_xcd(self._h, (C.c_int32 * L.NSTAGE)(*map(int, chunks))))

    @property
    def zmarch(self):
        """z planes each thread walks, per stage range (0: one point)."""
        t = (C.c_int32 * L.NSTAGE)()
        L.check(L.lib().camr_level_get_zmarch(self._h, t))
        return list(t)

    @zmarch.setter
    def zmarch(self, planes):
        L.check(L.lib().camr_level_set_zmarch(self._h, (C.c_int32 * L.NSTAGE)(*map(int, planes))))

    @property
    def ctu(self):
        """(on, planes): the LDS-staged CTU face kernels (1) or the staged
        transdo / transdd launches (0), and the z planes per CTU block."""
        on, planes = C.c_int32(), C.c_int32()
        L.check(L.lib().camr_level_get_ctu(self._h, C.byref(on), C.byref(planes)))
        return on.value, planes.value

    @ctu.setter
    def ctu(self, value):
        on, planes = value
        L.check(L.lib().camr_level_set_ctu(self._h, int(on), int(planes)))

    def profile(self, enable=True, kernel=None):
        """Time launches (all, or only those named `kernel`) with hipEvents."""
        L.check(L.lib().camr_level_profile(
            self._h, int(enable), kernel.encode() if kernel else None))

    def profile_report(self):
        import json
        buf = C.create_string_buffer(1 << 16)
        L.check(L.lib().camr_level_profile_report(self._h, buf, len(buf)))
        return json.loads(buf.value.decode())


# ---------------------------------------------------------------- per box
def array4_of(t, lo, ncomp=None) -> L.Array4:
    """Array4 view of a (nc, nz, ny, nx) contiguous float64 torch tensor."""
    assert t.dtype.is_floating_point and t.is_contiguous() and t.dim() == 4
    nc, nz, ny, nx = t.shape
    a = L.Array4()
    a.p = t.data_ptr()
    a.jstride, a.kstride, a.nstride = nx, nx * ny, nx * ny * nz
    for d in range(3):
        a.lo[d] = int(lo[d])
    a.hi[0], a.hi[1], a.hi[2] = lo[0] + nx - 1, lo[1] + ny - 1, lo[2] + nz - 1
    a.ncomp = nc if ncomp is None else ncomp
    return a


def _box(lo, hi) -> L.Box:
    b = L.Box()
    for d in range(3):
        b.lo[d], b.hi[d] = int(lo[d]), int(hi[d])
    return b


def _stream(stream):
    return C.c_void_p(stream) if stream else None


def hydro_ctoprim(qlo, qhi, u, ulo, q, qa, params, stream=0):
    L.check(L.lib().camr_hydro_ctoprim(
        C.byref(_box(qlo, qhi)), C.byref(array4_of(u, ulo)),
        C.byref(array4_of(q, qlo)), C.byref(array4_of(qa, qlo)),
        C.byref(params), _stream(stream)))


def hydro_srctoprim(qlo, qhi, q, qa, src, srcq, stream=0):
    L.check(L.lib().camr_hydro_srctoprim(
        C.byref(_box(qlo, qhi)), C.byref(array4_of(q, qlo)),
        C.byref(array4_of(qa, qlo)), C.byref(array4_of(src, qlo)),
        C.byref(array4_of(srcq, qlo)), _stream(stream)))


def scratch_bytes(lo, hi, params):
    """Bytes of scratch one hydro_umdrv call on box lo..hi takes."""
    n = C.c_size_t()
    L.check(L.lib().camr_scratch_bytes(C.byref(_box(lo, hi)), C.byref(params), C.byref(n)))
    return n.value


def hydro_umdrv(lo, hi, uin, uin_lo, dsdt, q, qa, srcq, flx, dt, params,
                area=None, vol=None, stream=0, noalloc=False):
    """hydro_umdrv (Hydro.H:11-35) on device tensors; q/qa/srcq start at
    lo-4, flx[d] at lo, dsdt at lo.  srcq None = identically +0 sources.
    noalloc: the allocation-free entry point (static device scratch)."""
    g = [x - L.NGROW for x in lo]
    fl = (L.Array4 * 3)(*[array4_of(f, lo) for f in flx])
    ar = (L.Array4 * 3)(*([array4_of(a, lo) for a in area] if area else
                          [L.Array4() for _ in range(3)]))
    fn = L.lib().camr_hydro_umdrv_noalloc if noalloc else L.lib().camr_hydro_umdrv
    L.check(fn(
        C.byref(_box(lo, hi)), C.byref(array4_of(uin, uin_lo)),
        C.byref(array4_of(dsdt, lo)), C.byref(array4_of(q, g)),
        C.byref(array4_of(qa, g)),
        C.byref(array4_of(srcq, g)) if srcq is not None else C.byref(L.Array4()),
        fl, ar, C.byref(array4_of(vol, lo)) if vol is not None else C.byref(L.Array4()),
        dt, C.byref(params), _stream(stream)))


def hydro_umdrv_eb(lo, hi, uin, dsdt, q, qa, flx, geom, dt, params, is_periodic,
                   stream=0):
    """hydro_umdrv_eb (Hydro.H:56-91, Godunov + FluxRedist) on device
    tensors: uin / q / qa start at lo-6, dsdt and flx[d] at lo.  `geom` is
    a dict of device tensors: flags (int32, from geom['flo']), vfrac, apx..,
    fcx.. (from lo-7)."""
    g6 = [x - 6 for x in lo]
    g7 = [x - 7 for x in lo]
    fl = (L.Array4 * 3)(*[array4_of(f, lo) for f in flx])
    G = L.EbGeom()
    G.flags = geom["flags"].data_ptr()
    fs = geom["flags"].shape
    for d in range(3):
        G.flo[d] = int(geom["flo"][d])
        G.fhi[d] = int(geom["flo"][d]) + fs[2 - d] - 1
    G.vfrac = array4_of(geom["vfrac"][None], g7)
    for d, n in enumerate("xyz"):
        G.ap[d] = array4_of(geom["ap" + n][None], g7)
        G.fc[d] = array4_of(geom["fc" + n], g7)
    per = (C.c_int32 * 3)(*[int(x) for x in is_periodic])
    L.check(L.lib().camr_hydro_umdrv_eb(
        C.byref(_box(lo, hi)), C.byref(array4_of(uin, g6)), C.byref(array4_of(dsdt, lo)),
        C.byref(array4_of(q, g6)), C.byref(array4_of(qa, g6)), C.byref(L.Array4()), fl,
        C.byref(G), dt, C.byref(params), per, _stream(stream)))


def estdt_hydro(lo, hi, u, ulo, dx, stream=0):
    out = C.c_double()
    dxa = (C.c_double * 3)(*dx)
    L.check(L.lib().camr_estdt_hydro(C.byref(_box(lo, hi)), C.byref(array4_of(u, ulo)),
                                     dxa, C.byref(out), _stream(stream)))
    return out.value


class FluxRegister:
    """The flux register between `coarse` and `fine` (AMReX's YAFluxRegister
    as CAMR uses it); attaching it makes both levels add their fluxes on
    every step.  Destroy it (close) before its levels."""

    def __init__(self, coarse: Level, fine: Level, ratio: int):
        self.coarse, self.fine, self.ratio = coarse, fine, int(ratio)
        h = C.c_void_p()
        L.check(L.lib().camr_flux_register_create(coarse._h, fine._h, self.ratio, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            L.lib().camr_flux_register_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def amr_step(coarse: Level, fine: Level, reg: FluxRegister, dt: float):
    """One subcycled two-level step of coarse size dt (Amr::timeStep with
    CAMR's post_timestep: reflux, avgDown, clean_state)."""
    L.check(L.lib().camr_amr_step(coarse._h, fine._h, reg._h, float(dt)))


def amr_advance(levels, regs, dt: float):
    """One coarse step of a hierarchy (levels coarse first, regs[l] joining
    levels l and l+1) with subcycling at every level."""
    n = len(levels)
    hl = (C.c_void_p * n)(*[lv._h.value for lv in levels])
    hr = (C.c_void_p * max(n - 1, 1))(*([r._h.value for r in regs] or [None]))
    L.check(L.lib().camr_amr_advance(hl, hr, n, float(dt)))
