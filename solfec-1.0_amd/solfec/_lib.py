"""ctypes bindings of the in-tree native libraries.

``lib/libsfgs.so`` is the HIP contact-solve library (``include/sfgs.h``) and
``lib/libsolfec.so`` the host C layer (``include/solfec.h``).  Both are built
by ``make -C solfec-1.0_amd`` (or ``__graft_entry__.build()``).  Loading fails
loudly when they are missing: there is no CPU fallback for the contact solve.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(os.path.dirname(_HERE), "lib")

_i = C.c_int
_d = C.c_double
_ll = C.c_longlong
_p = C.c_void_p
_s = C.c_char_p


class SFGS_PARAM(C.Structure):
    _fields_ = [("epsilon", _d), ("maxiter", _i), ("meritval", _d), ("failure", _i),
                ("diagepsilon", _d), ("diagmaxiter", _i), ("diagsolver", _i), ("reverse", _i),
                ("precision", _i), ("refine", _i), ("profile", _i)]


class SFGS_REPORT(C.Structure):
    _fields_ = [("iters", _i), ("error", _i), ("rel_error", _d), ("merit", _d), ("ncolours", _i),
                ("fail_rows", _i), ("aborted", _i), ("solve_ms", _d), ("sweep_ms", _d),
                ("sweep_launches", _ll), ("row_updates", _ll), ("nnz", _ll), ("fp64_sweeps", _i),
                ("launches_per_sweep", _i)]


class SFGS_NEWTON_PARAM(C.Structure):
    _fields_ = [("meritval", _d), ("maxiter", _i), ("linver", _i), ("linmaxiter", _i), ("maxmatvec", _i),
                ("epsilon", _d), ("delta", _d), ("theta", _d), ("omega", _d), ("gsflag", _i), ("reldelta", _i),
                ("bodyspace", _i)]


class SFGS_NEWTON_REPORT(C.Structure):
    _fields_ = [("iters", _i), ("matvecs", _i), ("converged", _i), ("gs_fallback", _i), ("gs_iters", _i),
                ("singular", _i), ("merit", _d), ("W_norm", _d)]


class SFGS_BATCH_TUNE(C.Structure):
    _fields_ = [("lanes_per_row", _i), ("waves_per_group", _i), ("rows_per_group", _i), ("schedule", _i)]


class SFGS_SCENE_REPORT(C.Structure):
    _fields_ = [("rows", _i), ("groups", _i), ("iters", _i), ("error_code", _i), ("fail_rows", _i),
                ("aborted", _i), ("error", _d), ("merit", _d), ("energy", _d)]


_SFGS_PROTOS = {
    "SFGS_Create": (_p, [_i]),
    "SFGS_Destroy": (None, [_p]),
    "SFGS_Error_String": (_s, [_i]),
    "SFGS_Last_Error": (_s, [_p]),
    "SFGS_Bodies_Upload": (_i, [_p, _i, _p, _p, _p, _p, _p, _p, _p]),
    "SFGS_Contacts_Upload": (_i, [_p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "SFGS_Locdyn_Update_Begin": (_i, [_p, _d, _i]),
    "SFGS_Locdyn_Update_Min": (_i, [_p, _d, _i]),
    "SFGS_System_Upload": (_i, [_p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _d, _d, _i, _p, _p]),
    "SFGS_Solve": (_i, [_p, C.POINTER(SFGS_PARAM), _p, _p, C.POINTER(SFGS_REPORT)]),
    "SFGS_Merit": (_i, [_p, _i, _p, _p]),
    "SFGS_Get_RU": (_i, [_p, _p, _p]),
    "SFGS_Update_Cohesion": (_i, [_p]),
    "SFGS_Get_Cohesion": (_i, [_p, _p]),
    "SFGS_Set_Cohesion": (_i, [_p, _p]),
    "SFGS_Bodyspace_Product": (_i, [_p, _p, _p]),
    "SFGS_Get_Locdyn": (_i, [_p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "SFGS_Get_Sizes": (_i, [_p, _p, _p, _p]),
    "SFGS_Colour": (_i, [_p, _i, _p, _p, _p]),
    "SFGS_Colour_Static": (_i, [_p, _i, C.c_longlong, _p, _p, _p]),
    "SFGS_Body_Update": (_i, [_p, _d, _p, _p]),
    "SFGS_Penalty_Solve": (_i, [_p, _d, _i, _p, _p, _p]),
    "SFGS_Newton_Solve": (_i, [_p, C.POINTER(SFGS_NEWTON_PARAM), _p, _p, C.POINTER(SFGS_NEWTON_REPORT)]),
    "SFGS_Set_Layout": (_i, [_p, _i]),
    "SFGS_Set_Schedule": (_i, [_p, _i]),
    "SFGS_Set_Sweep_Block": (_i, [_p, _i]),
    "SFGS_Set_Row_Order": (_i, [_p, _i]),
    "SFGS_Mlp_Set": (_i, [_p, _p, _ll, _p]),
    "SFGS_Mlp_Get": (_i, [_p, _p, _p]),
    "SFGS_Mlp_Forward": (_i, [_p, _p, _i, _p]),
    "SFGS_Mlp_Grad": (_i, [_p, _p, _p, _i, _p, _p]),
    "SFGS_Mlp_Train": (_i, [_p, _p, _p, _ll, _i, _i, C.c_float, C.c_ulonglong, _p]),
    "SFGS_Features": (_i, [_p, _p]),
    "SFGS_Warm_Start": (_i, [_p]),
    "SFGS_Warm_Start_Mode": (_i, [_p, _i, C.c_float]),
    "SFGS_Batch_Scenes": (_i, [_p, _i, _p]),
    "SFGS_Batch_Solve": (_i, [_p, C.POINTER(SFGS_PARAM), _p, _p, C.POINTER(SFGS_REPORT)]),
    "SFGS_Batch_Autotune": (_i, [_p, C.POINTER(SFGS_PARAM), _i, C.POINTER(SFGS_BATCH_TUNE), C.POINTER(_d)]),
}

_SOL_PROTOS = {
    "GAUSS_SEIDEL_Create": (_p, [_d, _i, _d, _i, _d, _i, _i, _p, _p]),
    "GAUSS_SEIDEL_Destroy": (None, [_p]),
    "GAUSS_SEIDEL_Get": (_d, [_p, _s]),
    "GAUSS_SEIDEL_Set": (_i, [_p, _s, _d]),
    "GAUSS_SEIDEL_Failure": (_s, [_p]),
    "GAUSS_SEIDEL_Diagsolver": (_s, [_p]),
    "GAUSS_SEIDEL_Error": (_s, [_p]),
    "GAUSS_SEIDEL_Reverse": (_s, [_p]),
    "GAUSS_SEIDEL_Variant": (_s, [_p]),
    "GAUSS_SEIDEL_History": (_i, [_p, _p, _p]),
    "GAUSS_SEIDEL_Itershist": (_i, [_p, _p, _i]),
    "GAUSS_SEIDEL_Report": (_i, [_p, C.POINTER(SFGS_REPORT)]),
    "SOLFEC_Create": (_p, [_i, _d, _s]),
    "SOLFEC_Destroy": (None, [_p]),
    "SOLFEC_Error_String": (_s, [_i]),
    "SOLFEC_Last_Error": (_s, [_p]),
    "SOLFEC_Gravity": (None, [_p, _d, _d, _d]),
    "SOLFEC_Surface_Material": (_i, [_p, _i, _i, _d, _d, _d]),
    "SOLFEC_Surface_Material_Ex": (_i, [_p, _i, _i, _i, _d, _d, _d, _d, _d, _d]),
    "NEWTON_Create": (_p, [_d, _i]),
    "NEWTON_Destroy": (None, [_p]),
    "NEWTON_Get": (_d, [_p, _s]),
    "NEWTON_Set": (_i, [_p, _s, _d]),
    "NEWTON_History": (_i, [_p, _p, _p]),
    "NEWTON_Itershist": (_i, [_p, _p, _i]),
    "NEWTON_Report": (_i, [_p, C.POINTER(SFGS_NEWTON_REPORT)]),
    "SOLFEC_Step_Solve_Newton": (_i, [_p, _p]),
    "SOLFEC_Run_Newton": (_i, [_p, _p, _d]),
    "PENALTY_Create": (_p, [_i]),
    "PENALTY_Destroy": (None, [_p]),
    "SOLFEC_Step_Solve_Penalty": (_i, [_p, _p]),
    "SOLFEC_Run_Penalty": (_i, [_p, _p, _d]),
    "SOLFEC_Sphere": (_i, [_p, _p, _d, _d, _i]),
    "SOLFEC_Plane": (_i, [_p, _p, _p, _i]),
    "SOLFEC_Set_Velocity": (_i, [_p, _i, _p]),
    "SOLFEC_Lattice_Scene": (_i, [_p, _i, _d, _d, _d, _d, _d, C.c_ulonglong]),
    "SOLFEC_Ellipsoid_Lattice_Scene": (_i, [_p, _i, _d, _d, _d, _d, _d, C.c_ulonglong]),
    "SOLFEC_Ellipsoid": (_i, [_p, _p, _p, _p, _d, _i]),
    "SOLFEC_Locdyn_Dump": (_i, [_p, C.c_char_p]),
    "SOLFEC_W_MatrixMarket": (_i, [_p, C.c_char_p]),
    "SOLFEC_Locdyn_Write": (_i, [C.c_char_p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "SOLFEC_W_MatrixMarket_Write": (_i, [C.c_char_p, _i, _p, _p, _p, _p]),
    "SOLFEC_Run": (_i, [_p, _p, _d]),
    "SOLFEC_Device_Loop": (_i, [_p, _i]),
    "SOLFEC_Write_State": (_i, [_p, C.c_char_p]),
    "SOLFEC_Read_State": (_i, [_p, C.c_char_p]),
    "SOLFEC_Step_Begin": (_i, [_p]),
    "SOLFEC_Step_Solve": (_i, [_p, _p]),
    "SOLFEC_Step_End": (_i, [_p]),
    "SOLFEC_Time": (_d, [_p]),
    "SOLFEC_Nbod": (_i, [_p]),
    "SOLFEC_Ncon": (_i, [_p]),
    "SOLFEC_Timing": (_d, [_p, _s]),
    "SOLFEC_Gpu": (_p, [_p]),
    "SOLFEC_Get_Bodies": (_i, [_p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "SOLFEC_Get_Contacts": (_i, [_p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "SOLFEC_Set_RU": (_i, [_p, _p, _p]),
    "SOLFEC_Capture": (_i, [_p, _p, _i]),
    "SOLFEC_Captured": (_i, [_p]),
}

_cache = {}


def _load(name, protos):
    if name in _cache:
        return _cache[name]
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `make -C solfec-1.0_amd` (or __graft_entry__.build()); "
                           "the contact solve has no CPU fallback")
    lib = C.CDLL(path)
    for fn, (res, args) in protos.items():
        f = getattr(lib, fn)
        f.restype = res
        f.argtypes = args
    _cache[name] = lib
    return lib


def sfgs():
    """The HIP contact-solve library (libsfgs.so)."""
    return _load("libsfgs.so", _SFGS_PROTOS)


def solfec_lib():
    """The host C layer (libsolfec.so); loads libsfgs.so first."""
    sfgs()
    return _load("libsolfec.so", _SOL_PROTOS)


def ptr(a):
    """Raw data pointer of a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the native layer must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None:
        assert a.size == n, (a.size, n)
    return a


def i32(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.int32)
    if n is not None:
        assert a.size == n, (a.size, n)
    return a
