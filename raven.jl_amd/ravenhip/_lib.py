"""ctypes binding of libravenhip.so (include/ravenhip.h).

The shared library must have been built (``__graft_entry__.build()`` or
``make -C raven.jl_amd/csrc``); importing the package without it raises, so a
missing HIP extension can never fall back to anything else.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libravenhip.so")

ABI_VERSION = 2
RV_OK = 0
RV_LAW_ADVECTION, RV_LAW_EULER = 1, 2
RV_FLUX_RUSANOV, RV_FLUX_CENTRAL, RV_FLUX_EC = 1, 2, 3
RV_FORM_WEAK, RV_FORM_FLUX_DIFFERENCING = 0, 1
RV_F64, RV_F32 = 1, 2
RV_GEOM_AUTO, RV_GEOM_AFFINE, RV_GEOM_GENERAL = 0, 1, 2
RV_LSRK54, RV_LSRK144 = 54, 144
RV_E_ARG, RV_E_UNSUPPORTED = -1, -4
# device scratch of rv_lsrk_run_noalloc (RV_NOALLOC_SCRATCH_BYTES)
RV_NOALLOC_SCRATCH_BYTES = 32 << 20

ERRORS = {-1: "RV_E_ARG", -2: "RV_E_HIP", -3: "RV_E_NOMEM", -4: "RV_E_UNSUPPORTED",
          -5: "RV_E_MAPS", -6: "RV_E_STATE"}

c_i32p = ctypes.POINTER(ctypes.c_int32)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_f64p = ctypes.POINTER(ctypes.c_double)
c_i64pp = ctypes.POINTER(ctypes.POINTER(ctypes.c_int64))
c_i8pp = ctypes.POINTER(ctypes.POINTER(ctypes.c_int8))


class RvDesc(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32), ("dim", ctypes.c_int32), ("nq", ctypes.c_int32),
        ("law", ctypes.c_int32), ("flux", ctypes.c_int32), ("dtype", ctypes.c_int32),
        ("geometry", ctypes.c_int32), ("directions", ctypes.c_int32),
        ("ncells", ctypes.c_int64), ("gamma", ctypes.c_double),
        ("velocity", ctypes.c_double * 3),
        ("D", c_f64p), ("weights", c_f64p), ("metrics", c_f64p),
        ("surface_metrics", c_f64p), ("vmapM", c_i64p), ("vmapP", c_i64p),
        ("boundary_codes", c_i64p), ("dtoc", c_i64p), ("ncontinuous", ctypes.c_int64),
        ("stream", ctypes.c_void_p),
        ("volume_form", ctypes.c_int32), ("volume_flux", ctypes.c_int32),
    ]


class RvBrickDesc(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32), ("dim", ctypes.c_int32), ("nq", ctypes.c_int32),
        ("law", ctypes.c_int32), ("flux", ctypes.c_int32), ("dtype", ctypes.c_int32),
        ("n", ctypes.c_int32 * 3), ("periodic", ctypes.c_int32 * 3), ("level", ctypes.c_int32),
        ("lo", ctypes.c_double * 3), ("hi", ctypes.c_double * 3), ("gamma", ctypes.c_double),
        ("velocity", ctypes.c_double * 3), ("stream", ctypes.c_void_p),
        ("volume_form", ctypes.c_int32), ("volume_flux", ctypes.c_int32),
    ]


class RvInfo(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32), ("nq", ctypes.c_int32), ("ns", ctypes.c_int32),
        ("dtype", ctypes.c_int32), ("geometry", ctypes.c_int32), ("faces_self", ctypes.c_int32),
        ("faces_explicit", ctypes.c_int32), ("streaming", ctypes.c_int32),
        ("ncells", ctypes.c_int64), ("state_bytes", ctypes.c_int64),
        ("device_bytes", ctypes.c_int64), ("stage_bytes", ctypes.c_int64),
        ("stage_bytes_first", ctypes.c_int64), ("rhs_bytes", ctypes.c_int64),
    ]


class RavenHIPError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_LIB = None
H = ctypes.c_void_p


def _hip_runtimes():
    try:
        with open("/proc/self/maps") as fh:
            return sorted({ln.split()[-1] for ln in fh if "libamdhip64" in ln})
    except OSError:
        return []


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"{LIBPATH} is missing: build it first (__graft_entry__.build())")
    # One process must hold one HIP runtime.  torch (which provides device
    # memory and streams here) ships its own libamdhip64 with the same soname;
    # loading it first makes libravenhip's dependency resolve to that copy.
    try:
        import torch  # noqa: F401
        torch.cuda.is_available()
    except ImportError:
        pass
    L = ctypes.CDLL(LIBPATH)
    runtimes = _hip_runtimes()
    if len(runtimes) > 1:
        raise RuntimeError("two HIP runtimes are loaded in this process: " + ", ".join(runtimes))
    vp = ctypes.c_void_p
    sig = {
        "rv_create": (ctypes.c_int, [ctypes.POINTER(RvDesc), ctypes.POINTER(H)]),
        "rv_create_brick": (ctypes.c_int, [ctypes.POINTER(RvBrickDesc), ctypes.POINTER(H)]),
        "rv_destroy": (None, [H]),
        "rv_last_error": (ctypes.c_char_p, [H]),
        "rv_get_info": (ctypes.c_int, [H, ctypes.POINTER(RvInfo)]),
        "rv_synchronize": (ctypes.c_int, [H]),
        "rv_stream": (vp, [H]),
        "rv_state_alloc": (ctypes.c_int, [H, ctypes.POINTER(vp)]),
        "rv_state_free": (ctypes.c_int, [H, vp]),
        "rv_state_upload": (ctypes.c_int, [H, vp, vp]),
        "rv_state_download": (ctypes.c_int, [H, vp, vp]),
        "rv_set_waves_per_block": (ctypes.c_int, [H, ctypes.c_int]),
        "rv_get_waves_per_block": (ctypes.c_int, [H]),
        "rv_rhs": (ctypes.c_int, [H, vp, vp, ctypes.c_double, ctypes.c_int]),
        "rv_lsrk_stage": (ctypes.c_int, [H, vp, vp, vp, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double]),
        "rv_lsrk_run": (ctypes.c_int, [H, vp, vp, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, c_f64p, c_i64p]),
        "rv_lsrk_run_noalloc": (ctypes.c_int, [H, vp, vp, ctypes.c_int, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_double, c_f64p,
                                               c_i64p]),
        "rv_lsrk_solve": (ctypes.c_int, [H, vp, vp, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         c_f64p, c_i64p]),
        "rv_weightednorm": (ctypes.c_int, [H, vp, c_f64p]),
        "rv_integrals": (ctypes.c_int, [H, vp, c_f64p]),
        "rv_weightednorm_p": (ctypes.c_int, [H, vp, ctypes.c_double, c_f64p]),
        "rv_entropy_integral": (ctypes.c_int, [H, vp, c_f64p]),
        "rv_entropy_product": (ctypes.c_int, [H, vp, vp, c_f64p]),
        "rv_rlsrk_step": (ctypes.c_int, [H, vp, vp, vp, vp, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, c_f64p,
                                         ctypes.POINTER(ctypes.c_int32)]),
        "rv_gather": (ctypes.c_int, [H, vp, vp, ctypes.c_int]),
        "rv_scatter": (ctypes.c_int, [H, vp, vp, ctypes.c_int]),
        "rv_get_maps": (ctypes.c_int, [H, c_i64p, c_i64p, c_i64p, c_i64p]),
        "rv_brick_numcells": (ctypes.c_int64, [ctypes.POINTER(RvBrickDesc)]),
        "rv_brick_arrays": (ctypes.c_int, [ctypes.POINTER(RvBrickDesc), c_f64p, c_f64p, c_f64p,
                                           c_i64p, c_i64p, c_i64p, c_i64p, c_i64p]),
        "rv_lobatto": (ctypes.c_int, [ctypes.c_int, c_f64p, c_f64p, c_f64p]),
        "rv_brick_cells": (ctypes.c_int, [ctypes.POINTER(RvBrickDesc), c_i32p]),
        "rv_forest_create": (ctypes.c_int, [ctypes.POINTER(RvBrickDesc), ctypes.POINTER(H)]),
        "rv_forest_destroy": (None, [H]),
        "rv_forest_numcells": (ctypes.c_int64, [H]),
        "rv_forest_adapt": (ctypes.c_int, [H, ctypes.POINTER(ctypes.c_uint8)]),
        "rv_forest_cells": (ctypes.c_int, [H, c_i32p, c_i32p, c_i32p]),
        "rv_forest_arrays": (ctypes.c_int, [H, ctypes.c_int32, c_f64p, c_i64p, c_i64p, c_i64p,
                                            c_i32p]),
        "rv_forest_mortars": (ctypes.c_int, [H, c_i64p]),
        "rv_forest_nc_counts": (ctypes.c_int, [H, c_i64p, c_i64p]),
        "rv_forest_nctables": (ctypes.c_int, [H, ctypes.c_int32, c_i64pp, c_i8pp, c_i8pp,
                                              c_i64pp]),
        "rv_mortars_from_nctables": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32,
                                                    ctypes.c_int64, c_i64p, c_i64pp, c_i8pp,
                                                    c_i32p, c_i64p]),
        "rv_forest_dtoc": (ctypes.c_int, [H, ctypes.c_int32, c_i64p, c_i64p]),
        "rv_create_forest": (ctypes.c_int, [H, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.POINTER(H)]),
        "rv_lobatto_halves": (ctypes.c_int, [ctypes.c_int, c_f64p, c_f64p]),
        "rv_enable_mortars": (ctypes.c_int, [H, c_i32p, c_i64p]),
        "rv_min_node_distance": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int64, vp,
                                                ctypes.c_int32, c_f64p, vp]),
        "rv_curved_metrics": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int64, vp,
                                             c_f64p, c_f64p, vp, vp, vp]),
        "rv_lsrk_coefficients": (ctypes.c_int, [ctypes.c_int, c_f64p, c_f64p, c_f64p,
                                                ctypes.POINTER(ctypes.c_int)]),
        "rv_batch_create": (ctypes.c_int, [ctypes.POINTER(H), ctypes.c_int32,
                                           ctypes.POINTER(H)]),
        "rv_batch_destroy": (None, [H]),
        "rv_batch_last_error": (ctypes.c_char_p, [H]),
        "rv_batch_info": (ctypes.c_int, [H, c_i64p, c_i64p, c_i32p]),
        "rv_batch_configure": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int64]),
        "rv_batch_set_streams": (ctypes.c_int, [H, ctypes.c_int32]),
        "rv_batch_stage": (ctypes.c_int, [H, ctypes.POINTER(vp), ctypes.POINTER(vp),
                                          ctypes.POINTER(vp), c_f64p, c_f64p]),
        "rv_batch_step": (ctypes.c_int, [H, ctypes.POINTER(vp), ctypes.POINTER(vp),
                                         ctypes.POINTER(vp), c_i32p, c_f64p, ctypes.c_int]),
        "rv_batch_steps": (ctypes.c_int, [H, ctypes.POINTER(vp), ctypes.POINTER(vp),
                                          ctypes.POINTER(vp), c_i32p, c_f64p, ctypes.c_int,
                                          ctypes.c_int64]),
        "rv_batch_lsrk_run": (ctypes.c_int, [H, ctypes.POINTER(vp), ctypes.POINTER(vp),
                                             ctypes.c_int, c_f64p, ctypes.c_int64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def check_batch(rc, b):
    if rc != RV_OK:
        msg = lib().rv_batch_last_error(b)
        raise RavenHIPError(rc, msg.decode() if msg else "")
    return rc


def check(rc, h=None):
    if rc != RV_OK:
        msg = lib().rv_last_error(h)
        raise RavenHIPError(rc, msg.decode() if msg else "")
    return rc


def ptr(a, ctype):
    """ctypes pointer to a numpy array (None passes NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))
