"""ctypes binding of libmfvm.so (the C-ABI declared in include/mfvm.h).

The library is required: importing this module on a machine without the
built library raises immediately, and there is no other compute path.
"""

import ctypes
import os

# torch first: its HIP runtime is then the one libmfvm.so binds to (one
# runtime per process, so device pointers and streams are shared with torch).
import torch  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MFVM_LIB") or os.path.join(_HERE, "lib", "libmfvm.so")

NMET = 19
(MET_V, MET_S1, MET_S2, MET_S3, MET_L1, MET_L2, MET_L3, MET_H1, MET_H2, MET_H3,
 MET_D1H, MET_D1H2, MET_D1H3, MET_D2H, MET_D2H1, MET_D2H3, MET_D3H, MET_D3H1,
 MET_D3H2) = range(NMET)

CART, CYL, SPH = 0, 1, 2
LIM_NONE, LIM_MINMOD, LIM_MC, LIM_VANLEER = 0, 1, 2, 3
RS_HLL, RS_RUSANOV = 0, 1
BC_NEUMANN, BC_PERIODIC, BC_REFLECTING = 0, 1, 2
INT_EULER, INT_RK3, INT_SSPRK3, INT_RK4 = 0, 1, 2, 3
F64, F32 = 0, 1
UPD_NONE, UPD_EULER, UPD_BASE, UPD_SSP, UPD_ACC6 = range(5)
ACC_NONE, ACC_SET, ACC_ADD = range(3)
OK, E_ARG, E_HIP, E_UNSUP = 0, -1, -2, -3


class Grid(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32 * 3), ("coord", ctypes.c_int32),
                ("fac", ctypes.c_void_p), ("fstride", ctypes.c_int32),
                ("uniform", ctypes.c_int32), ("lpow2", ctypes.c_int32 * 3),
                ("vpow2", ctypes.c_int32), ("lval", ctypes.c_double * 3),
                ("vval", ctypes.c_double), ("sval", ctypes.c_double * 3),
                ("smax", ctypes.c_double), ("lmin", ctypes.c_double * 3)]


class NSParams(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int32), ("limiter", ctypes.c_int32),
                ("riemann", ctypes.c_int32), ("bc", (ctypes.c_int32 * 6) * 5),
                ("gamma", ctypes.c_double), ("mu", ctypes.c_double),
                ("g", ctypes.c_double * 3), ("tile", ctypes.c_int32 * 3)]


class Stage(ctypes.Structure):
    _fields_ = [("U", ctypes.c_void_p), ("C", ctypes.c_void_p), ("U0", ctypes.c_void_p),
                ("out", ctypes.c_void_p), ("kacc", ctypes.c_void_p),
                ("update", ctypes.c_int32), ("acc", ctypes.c_int32),
                ("c0", ctypes.c_double), ("c1", ctypes.c_double),
                ("acc_coef", ctypes.c_double), ("dt", ctypes.c_double),
                ("dt_dev", ctypes.c_void_p), ("partials", ctypes.c_void_p),
                ("lo", ctypes.c_void_p), ("dinc", ctypes.c_void_p), ("snext", ctypes.c_void_p),
                ("comp", ctypes.c_int32), ("xy", ctypes.c_void_p)]


class NSCompWork(ctypes.Structure):
    _fields_ = [("hi", ctypes.c_void_p), ("lo", ctypes.c_void_p), ("Sa", ctypes.c_void_p),
                ("Sb", ctypes.c_void_p), ("C", ctypes.c_void_p), ("dinc", ctypes.c_void_p),
                ("kacc", ctypes.c_void_p), ("partials", ctypes.c_void_p), ("dts", ctypes.c_void_p),
                ("xy", ctypes.c_void_p)]


class HeatParams(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int32), ("limiter", ctypes.c_int32),
                ("bc", (ctypes.c_int32 * 6) * 1), ("rho_cp", ctypes.c_double),
                ("conductivity", ctypes.c_double), ("source", ctypes.c_double),
                ("source_field", ctypes.c_void_p)]


class MaxwellParams(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int32), ("limiter", ctypes.c_int32),
                ("riemann", ctypes.c_int32), ("bc", (ctypes.c_int32 * 6) * 6),
                ("mu", ctypes.c_double), ("ep", ctypes.c_double), ("sigma", ctypes.c_double),
                ("rho", ctypes.c_double), ("chi_b", ctypes.c_double), ("chi_d", ctypes.c_double),
                ("sigma_field", ctypes.c_void_p), ("rho_field", ctypes.c_void_p)]


class NSWork(ctypes.Structure):
    _fields_ = [("U", ctypes.c_void_p), ("S1", ctypes.c_void_p), ("S2", ctypes.c_void_p),
                ("C", ctypes.c_void_p), ("kacc", ctypes.c_void_p),
                ("partials", ctypes.c_void_p), ("dts", ctypes.c_void_p), ("xy", ctypes.c_void_p)]


ENS_ORDER_MEMBER, ENS_ORDER_COST, ENS_ORDER_INTERLEAVE = 0, 1, 2


class EnsMember(ctypes.Structure):
    _fields_ = [("grid", Grid), ("bc", (ctypes.c_int32 * 6) * 5), ("tile", ctypes.c_int32 * 3),
                ("U", ctypes.c_void_p), ("S1", ctypes.c_void_p), ("S2", ctypes.c_void_p),
                ("C", ctypes.c_void_p), ("kacc", ctypes.c_void_p), ("lo", ctypes.c_void_p),
                ("dinc", ctypes.c_void_p), ("partials", ctypes.c_void_p), ("dts", ctypes.c_void_p),
                ("xy", ctypes.c_void_p)]


class EnsOpts(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int32), ("persist", ctypes.c_int32), ("graph", ctypes.c_int32),
                ("remap", ctypes.c_int32)]


CNN_FWD_RELU, CNN_FWD_LINEAR, CNN_BWD_DATA = 0, 1, 2


class CNN(ctypes.Structure):
    _fields_ = [("wf", ctypes.c_void_p * 3), ("b", ctypes.c_void_p * 3), ("in_shift", ctypes.c_void_p),
                ("in_scale", ctypes.c_void_p), ("out_scale", ctypes.c_void_p)]


class ClosureWork(ctypes.Structure):
    _fields_ = [("x", ctypes.c_void_p), ("a1", ctypes.c_void_p), ("a2", ctypes.c_void_p),
                ("y", ctypes.c_void_p), ("T", ctypes.c_void_p), ("flag", ctypes.c_void_p),
                ("counts", ctypes.c_void_p), ("dt_cur", ctypes.c_void_p), ("t_cur", ctypes.c_void_p),
                ("step", ctypes.c_void_p), ("layered", ctypes.c_int32)]


# (name, restype, argtypes) of every exported entry point.
_P = ctypes.c_void_p
_G = ctypes.POINTER(Grid)
_NP = ctypes.POINTER(NSParams)
SIGNATURES = [
    ("mfvm_last_error", ctypes.c_char_p, []),
    ("mfvm_abi_version", ctypes.c_int32, []),
    ("mfvm_timing", ctypes.c_int, [ctypes.c_int32, _P, _P]),
    ("mfvm_apply_bcs", ctypes.c_int, [_G, ctypes.c_int32, _P, ctypes.c_int32, _P, _P, _P]),
    ("mfvm_copy_interior", ctypes.c_int, [_G, ctypes.c_int32, ctypes.c_int32, _P, _P, _P]),
    ("mfvm_ns_closure", ctypes.c_int, [_G, _NP, _P, _P, _P]),
    ("mfvm_ns_rhs", ctypes.c_int, [_G, _NP, _P, _P, _P, _P]),
    ("mfvm_ns_dt", ctypes.c_int, [_G, _NP, _P, ctypes.c_double, _P, _P, _P]),
    ("mfvm_ns_dt_noalloc", ctypes.c_int, [_G, _NP, _P, ctypes.c_double, _P, _P]),
    ("mfvm_ns_stage", ctypes.c_int, [_G, _NP, ctypes.POINTER(Stage), _P]),
    ("mfvm_ns_step", ctypes.c_int, [_G, _NP, ctypes.POINTER(NSWork), ctypes.c_int32,
                                    ctypes.c_double, _P]),
    ("mfvm_ns_advance", ctypes.c_int, [_G, _NP, ctypes.POINTER(NSWork), ctypes.c_int32,
                                       ctypes.c_double, ctypes.c_int32, _P]),
    ("mfvm_ns_partials_size", ctypes.c_int64, [_G, _NP]),
    ("mfvm_ns_comp_advance", ctypes.c_int, [_G, _NP, ctypes.POINTER(NSCompWork), ctypes.c_int32,
                                            ctypes.c_double, ctypes.c_int32, _P]),
    ("mfvm_ens_member_tiles", ctypes.c_int64, [ctypes.POINTER(EnsMember)]),
    ("mfvm_ens_table_bytes", ctypes.c_int64, [ctypes.POINTER(EnsMember), ctypes.c_int32,
                                              ctypes.c_int32]),
    ("mfvm_ens_advance", ctypes.c_int, [ctypes.POINTER(EnsMember), ctypes.c_int32, _NP,
                                        ctypes.POINTER(EnsOpts), ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_double, ctypes.c_int32, _P, ctypes.c_int64, _P]),
    ("mfvm_cnn_conv", ctypes.c_int, [ctypes.c_int32] * 6 + [_P, _P, _P, _P, _P, _P]),
    ("mfvm_cnn_wgrad", ctypes.c_int, [ctypes.c_int32] * 5 + [_P, _P, _P, _P, _P]),
    ("mfvm_cnn_pack", ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _P, _P, _P, _P]),
    ("mfvm_cnn_adam", ctypes.c_int, [ctypes.c_int64, _P, _P, _P, _P, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_int32, _P]),
    ("mfvm_cnn_features", ctypes.c_int, [_G, ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                         ctypes.c_int64, _P, _P, _P, _P, _P]),
    ("mfvm_cnn_loss", ctypes.c_int, [ctypes.c_int32] * 3 + [ctypes.c_double, ctypes.c_double,
                                                            _P, _P, _P, _P, _P, _P, _P]),
    ("mfvm_closure_advance", ctypes.c_int, [_G, _NP, ctypes.POINTER(NSWork), ctypes.POINTER(CNN),
                                            ctypes.POINTER(ClosureWork), ctypes.c_double,
                                            ctypes.c_double, ctypes.c_int32, ctypes.c_int32, _P]),
    ("mfvm_heat_rhs", ctypes.c_int, [_G, ctypes.POINTER(HeatParams), _P, _P, _P]),
    ("mfvm_heat_dt", ctypes.c_int, [_G, ctypes.POINTER(HeatParams), ctypes.c_double, _P, _P]),
    ("mfvm_heat_step", ctypes.c_int, [_G, ctypes.POINTER(HeatParams), ctypes.POINTER(NSWork),
                                      ctypes.c_int32, ctypes.c_double, _P]),
    ("mfvm_heat_advance", ctypes.c_int, [_G, ctypes.POINTER(HeatParams), ctypes.POINTER(NSWork),
                                         ctypes.c_int32, ctypes.c_double, ctypes.c_int32, _P]),
    ("mfvm_maxwell_residual", ctypes.c_int, [_G, ctypes.POINTER(MaxwellParams), _P, _P, _P]),
    ("mfvm_maxwell_rhs", ctypes.c_int, [_G, ctypes.POINTER(MaxwellParams), _P, _P, _P, _P]),
    ("mfvm_maxwell_dt", ctypes.c_int, [_G, ctypes.POINTER(MaxwellParams), ctypes.c_double, _P, _P]),
    ("mfvm_maxwell_step", ctypes.c_int, [_G, ctypes.POINTER(MaxwellParams), ctypes.POINTER(NSWork),
                                         ctypes.c_int32, ctypes.c_double, _P]),
    ("mfvm_maxwell_advance", ctypes.c_int, [_G, ctypes.POINTER(MaxwellParams),
                                            ctypes.POINTER(NSWork), ctypes.c_int32,
                                            ctypes.c_double, ctypes.c_int32, _P]),
    ("mfvm_diag_partials_size", ctypes.c_int64, [_G, ctypes.c_int32]),
    ("mfvm_diagnostics", ctypes.c_int, [_G, ctypes.c_int32, ctypes.c_int32, _P, _P, _P, ctypes.c_double,
                                        _P, _P, _P]),
]


class MfvmError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libmfvm.so not found at {LIB_PATH}; build it first "
            "(python -c 'import __graft_entry__ as g; g.build()').")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(rc):
    """Map a C-ABI status to the reference's exception types."""
    if rc == 0:
        return
    msg = lib.mfvm_last_error().decode()
    if rc == E_ARG:
        raise ValueError(msg)
    if rc == E_UNSUP:
        raise NotImplementedError(msg)
    raise MfvmError(msg)


TK_STAGE, TK_CLOSURE, TK_BC, TK_OTHER = range(4)


def timing(mode):
    """mfvm_timing wrapper: mode 1 start, 0 stop, 2 -> (ms[4], count[4])."""
    ms = (ctypes.c_double * 4)()
    cnt = (ctypes.c_int64 * 4)()
    check(lib.mfvm_timing(mode, ms, cnt))
    if mode == 2:
        return list(ms), list(cnt)
    return None


def stream_ptr(torch_mod):
    """Raw hipStream_t of torch's current stream."""
    return ctypes.c_void_p(torch_mod.cuda.current_stream().cuda_stream)
