"""ctypes binding of ``libcamr_hydro.so`` (the C-ABI in include/camr_hydro.h).

The shared library is built in-tree by ``camr_amd/build.py`` (or
``__graft_entry__.build()``).  Loading fails loudly when it is missing: there
is no CPU fallback for the hydro path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CAMR_LIB_PATH") or os.path.join(HERE, "libcamr_hydro.so")

NVAR, QVAR, NQAUX, NGROW = 10, 11, 5, 4

# status codes (camr_status)
OK, ERR_PARAM, ERR_SHAPE, ERR_NEG_ENERGY, ERR_ZERO_DENSITY, ERR_DT, ERR_HIP, ERR_ALLOC = range(8)

# physical BC codes (camr_bc)
INTERIOR, INFLOW, OUTFLOW, SYMMETRY, SLIPWALL, NOSLIPWALL = range(6)

ENGINE_AUTO, ENGINE_STAGED, ENGINE_FUSED = 0, 1, 2


class Box(C.Structure):
    _fields_ = [("lo", C.c_int32 * 3), ("hi", C.c_int32 * 3)]


class Array4(C.Structure):
    _fields_ = [("p", C.c_void_p), ("jstride", C.c_int64),
                ("kstride", C.c_int64), ("nstride", C.c_int64),
                ("lo", C.c_int32 * 3), ("hi", C.c_int32 * 3),
                ("ncomp", C.c_int32)]


class EbGeom(C.Structure):
    _fields_ = [("flags", C.c_void_p), ("flo", C.c_int32 * 3), ("fhi", C.c_int32 * 3),
                ("vfrac", Array4), ("ap", Array4 * 3), ("fc", Array4 * 3)]


class PlotDesc(C.Structure):
    _fields_ = [("nlev", C.c_int32), ("levels", C.POINTER(C.c_void_p)),
                ("ref_ratio", C.POINTER(C.c_int32)), ("level_steps", C.POINTER(C.c_int32)),
                ("prob_lo", C.c_double * 3), ("prob_hi", C.c_double * 3),
                ("plot_rhoy", C.c_int32), ("derive_vars", C.c_char_p)]


class CheckpointDesc(C.Structure):
    _fields_ = [("nlev", C.c_int32), ("levels", C.POINTER(C.c_void_p)),
                ("ref_ratio", C.POINTER(C.c_int32)), ("level_steps", C.POINTER(C.c_int32)),
                ("dt_level", C.POINTER(C.c_double))]


class HydroParams(C.Structure):
    _fields_ = [
        ("ppm_type", C.c_int32), ("plm_iorder", C.c_int32),
        ("use_pslope", C.c_int32), ("use_flattening", C.c_int32),
        ("transverse_reset_density", C.c_int32),
        ("allow_negative_energy", C.c_int32),
        ("allow_small_energy", C.c_int32),
        ("dual_energy_update_E_from_e", C.c_int32),
        ("small", C.c_double), ("small_dens", C.c_double),
        ("small_pres", C.c_double), ("small_temp", C.c_double),
        ("smallu", C.c_double), ("small_num", C.c_double),
        ("difmag", C.c_double), ("dual_energy_eta1", C.c_double),
        ("dual_energy_eta2", C.c_double),
        ("bclo", C.c_int32 * 3), ("bchi", C.c_int32 * 3),
        ("domlo", C.c_int32 * 3), ("domhi", C.c_int32 * 3),
        ("dx", C.c_double * 3), ("do_mol", C.c_int32),
    ]


class LevelDesc(C.Structure):
    _fields_ = [
        ("hp", HydroParams), ("is_periodic", C.c_int32 * 3),
        ("add_grav_src", C.c_int32), ("const_grav", C.c_double),
        ("cfl", C.c_double), ("change_max", C.c_double),
        ("init_shrink", C.c_double), ("max_dt", C.c_double),
        ("fixed_dt", C.c_double), ("stop_time", C.c_double),
        ("nbox", C.c_int32), ("box_lohi", C.POINTER(C.c_int32)),
        ("engine", C.c_int32), ("batched", C.c_int32), ("graph", C.c_int32),
        ("worker_streams", C.c_int32),
    ]

NSTAGE = 13  # CAMR_NSTAGE: stage ranges of the fused engine


# Every extern "C" symbol declared in include/camr_hydro.h.
EXPORTS = [
    "camr_default_params", "camr_last_error", "camr_version",
    "camr_hydro_ctoprim", "camr_hydro_srctoprim", "camr_hydro_umdrv", "camr_scratch_bytes",
    "camr_hydro_umdrv_noalloc",
    "camr_estdt_hydro", "camr_default_level_desc", "camr_level_create",
    "camr_level_destroy", "camr_level_set_stream", "camr_level_upload_box",
    "camr_level_download_box", "camr_level_state_array",
    "camr_level_init_post", "camr_level_initial_dt", "camr_level_new_dt",
    "camr_level_advance", "camr_level_run", "camr_level_sync",
    "camr_level_time", "camr_level_profile", "camr_level_profile_report",
    "camr_level_set_exterior", "camr_level_set_tuning", "camr_level_get_tuning",
    "camr_level_set_tiles", "camr_level_get_tiles", "camr_level_set_xcd", "camr_level_get_xcd",
    "camr_level_set_zmarch", "camr_level_get_zmarch", "camr_level_set_ctu",
    "camr_level_get_ctu",
    "camr_level_adjoint_forward", "camr_level_adj_upload_box", "camr_level_adj_download_box",
    "camr_level_adjoint_backward", "camr_level_adj_seed_eb_force", "camr_hydro_umdrv_eb", "camr_level_set_eb",
    "camr_level_set_eb_merge", "camr_level_eb_merge_info", "camr_level_box_kind", "camr_level_set_box_eb", "camr_level_exterior_from_coarse",
    "camr_flux_register_create", "camr_flux_register_destroy", "camr_level_average_down",
    "camr_amr_step", "camr_write_plotfile", "camr_level_eb_force", "camr_amr_advance",
    "camr_write_checkpoint", "camr_level_set_time",
]

_lib = None


class CamrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"camr status {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is not built; run camr_amd/build.py "
                          "(there is no CPU fallback for the hydro path)")
        # one HIP runtime per process: when PyTorch is installed its bundled
        # runtime is loaded first, so that the library and torch's device
        # tensors share it
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_int
        L.camr_default_params.argtypes = [C.POINTER(HydroParams)]
        L.camr_default_params.restype = None
        L.camr_default_level_desc.argtypes = [C.POINTER(LevelDesc)]
        L.camr_default_level_desc.restype = None
        L.camr_last_error.restype = C.c_char_p
        L.camr_version.restype = ip
        A4 = C.POINTER(Array4)
        L.camr_hydro_ctoprim.argtypes = [C.POINTER(Box), A4, A4, A4, C.POINTER(HydroParams), vp]
        L.camr_hydro_srctoprim.argtypes = [C.POINTER(Box), A4, A4, A4, A4, vp]
        L.camr_hydro_umdrv.argtypes = [C.POINTER(Box), A4, A4, A4, A4, A4, A4, A4, A4,
                                       C.c_double, C.POINTER(HydroParams), vp]
        L.camr_hydro_umdrv_noalloc.argtypes = L.camr_hydro_umdrv.argtypes
        L.camr_scratch_bytes.argtypes = [C.POINTER(Box), C.POINTER(HydroParams),
                                         C.POINTER(C.c_size_t)]
        L.camr_estdt_hydro.argtypes = [C.POINTER(Box), A4, dp, dp, vp]
        L.camr_level_create.argtypes = [C.POINTER(LevelDesc), C.POINTER(vp)]
        L.camr_level_destroy.argtypes = [vp]
        L.camr_level_set_stream.argtypes = [vp, vp]
        L.camr_level_upload_box.argtypes = [vp, C.c_int32, dp]
        L.camr_level_download_box.argtypes = [vp, C.c_int32, dp]
        L.camr_level_state_array.argtypes = [vp, C.c_int32, A4]
        L.camr_level_init_post.argtypes = [vp]
        L.camr_level_initial_dt.argtypes = [vp, dp]
        L.camr_level_new_dt.argtypes = [vp, C.c_double, dp]
        L.camr_level_advance.argtypes = [vp, C.c_double]
        L.camr_level_run.argtypes = [vp, C.c_int32, dp]
        L.camr_level_sync.argtypes = [vp]
        L.camr_level_time.argtypes = [vp]
        L.camr_level_time.restype = C.c_double
        L.camr_level_profile.argtypes = [vp, C.c_int32, C.c_char_p]
        L.camr_level_profile_report.argtypes = [vp, C.c_char_p, C.c_int64]
        L.camr_level_set_exterior.argtypes = [vp, dp]
        L.camr_level_set_tuning.argtypes = [vp, C.POINTER(C.c_int32), C.c_int32]
        L.camr_level_get_tuning.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.camr_level_set_tiles.argtypes = [vp, C.POINTER(C.c_int32)]
        L.camr_level_get_tiles.argtypes = [vp, C.POINTER(C.c_int32)]
        L.camr_level_set_xcd.argtypes = [vp, C.POINTER(C.c_int32)]
        L.camr_level_get_xcd.argtypes = [vp, C.POINTER(C.c_int32)]
        L.camr_level_set_zmarch.argtypes = [vp, C.POINTER(C.c_int32)]
        L.camr_level_get_zmarch.argtypes = [vp, C.POINTER(C.c_int32)]
        L.camr_level_set_ctu.argtypes = [vp, C.c_int32, C.c_int32]
        L.camr_level_get_ctu.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.camr_level_adjoint_forward.argtypes = [vp, C.c_int32, dp, C.c_int32]
        L.camr_level_adj_upload_box.argtypes = [vp, C.c_int32, dp]
        L.camr_level_adj_download_box.argtypes = [vp, C.c_int32, dp]
        L.camr_level_adjoint_backward.argtypes = [vp]
        L.camr_level_adj_seed_eb_force.argtypes = [vp, dp]
        i32p = C.POINTER(C.c_int32)
        L.camr_level_set_eb.argtypes = [vp, i32p, i32p, i32p]
        L.camr_level_set_eb_merge.argtypes = [vp, C.c_int32]
        L.camr_level_eb_merge_info.argtypes = [vp, i32p, i32p]
        L.camr_level_exterior_from_coarse.argtypes = [vp, vp, C.c_int32]
        L.camr_flux_register_create.argtypes = [vp, vp, C.c_int32, C.POINTER(vp)]
        L.camr_flux_register_destroy.argtypes = [vp]
        L.camr_level_average_down.argtypes = [vp, vp, C.c_int32]
        L.camr_amr_step.argtypes = [vp, vp, vp, C.c_double]
        L.camr_amr_advance.argtypes = [C.POINTER(vp), C.POINTER(vp), C.c_int32, C.c_double]
        L.camr_write_plotfile.argtypes = [C.c_char_p, C.POINTER(PlotDesc)]
        L.camr_write_checkpoint.argtypes = [C.c_char_p, C.POINTER(CheckpointDesc)]
        L.camr_level_set_time.argtypes = [vp, C.c_double]
        L.camr_level_eb_force.argtypes = [vp, C.POINTER(C.c_double)]
        L.camr_hydro_umdrv_eb.argtypes = [C.POINTER(Box), A4, A4, A4, A4, A4, A4,
                                          C.POINTER(EbGeom), C.c_double,
                                          C.POINTER(HydroParams), i32p, vp]
        L.camr_level_box_kind.argtypes = [vp, C.c_int32, i32p]
        L.camr_level_set_box_eb.argtypes = [vp, C.c_int32] + [dp] * 7
        for name in EXPORTS:
            f = getattr(L, name)
            if f.restype is C.c_int:  # default restype
                pass
        _lib = L
    return _lib


def check(rc):
    if rc != OK:
        msg = lib().camr_last_error().decode(errors="replace")
        raise CamrError(rc, msg)
    return rc
