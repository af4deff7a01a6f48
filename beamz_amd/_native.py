"""ctypes binding of ``libbeamz_amd.so`` (declarations: include/beamz_amd.h).

The library is built in-tree by ``beamz_amd.build``. Loading never falls back
to anything else: a missing library raises ``RuntimeError`` naming the build
command, and every failing entry point raises with the library's own message.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbeamz_amd.so")

i32, i64, f32, f64, vp = C.c_int, C.c_int64, C.c_float, C.c_double, C.c_void_p
P6 = vp * 6


class Shape(C.Structure):
    _fields_ = [("kind", i32), ("perm", f64), ("permb", f64), ("cond", f64), ("bbox", f64 * 6),
                ("px", f64), ("py", f64), ("w", f64), ("h", f64), ("cx", f64), ("cy", f64), ("r", f64),
                ("r_in", f64), ("r_out", f64), ("nv", i32), ("v", C.POINTER(f64)), ("nholes", i32),
                ("hole_nv", C.POINTER(i32)), ("hv", C.POINTER(f64)), ("cell", C.POINTER(f64)),
                ("cell_i0", i32), ("cell_j0", i32), ("cell_ni", i32), ("cell_nj", i32)]


class Pml(C.Structure):
    _fields_ = [("kind", i32), ("orient", i32), ("px", f64), ("py", f64), ("w", f64), ("h", f64), ("r", f64),
                ("order", f64), ("sigma_factor", f64), ("alpha_max", f64)]


class BoxSource(C.Structure):
    _fields_ = [("comp", i32), ("x0", i32), ("y0", i32), ("z0", i32), ("sx", i32), ("sy", i32), ("sz", i32),
                ("add", vp), ("mul", vp), ("amp", f32)]


MAX_BOX = 4


class Sim2D(C.Structure):
    _fields_ = [("nx", i32), ("ny", i32), ("ld", i32), ("ez", vp * 2), ("hx", vp * 2), ("hy", vp * 2),
                ("eps", vp), ("sig", vp), ("kh", f32), ("kbh", f32), ("ke", f32), ("kbe", f32),
                ("rdx", f32), ("rdy", f32), ("src_x0", i32), ("src_y0", i32), ("src_sx", i32), ("src_sy", i32),
                ("src_add", vp), ("src_mul", vp), ("signal", vp), ("power", vp),
                ("snap_x0", i32), ("snap_y0", i32), ("snap_sx", i32), ("snap_sy", i32), ("snap", vp),
                ("stg", vp), ("stg_x0", i32), ("stg_y0", i32), ("stg_sx", i32), ("stg_sy", i32), ("coef", vp),
                ("src_sel", vp), ("src_sig_len", i32)]


class Adj2D(C.Structure):
    _fields_ = [("nx", i32), ("ny", i32), ("ld", i32), ("le", vp * 2), ("lhx", vp * 2), ("lhy", vp * 2),
                ("eps", vp), ("sig", vp), ("kh", f32), ("kbh", f32), ("ke", f32), ("kbe", f32),
                ("rdx", f32), ("rdy", f32), ("src_x0", i32), ("src_y0", i32), ("src_sx", i32), ("src_sy", i32),
                ("src_mul", vp), ("box_x0", i32), ("box_y0", i32), ("box_sx", i32), ("box_sy", i32),
                ("snap", vp), ("snap_step0", i32), ("grad", vp),
                ("inj_x0", i32), ("inj_y0", i32), ("inj_sx", i32), ("inj_sy", i32), ("inj_w", vp), ("phasor", vp), ("nf", i32),
                ("coef", vp), ("l2hist", vp), ("inj_tab", vp)]

# name -> argtypes (restype is int unless listed in _RESTYPE)
SIGNATURES = {
    "bz_last_error": [],
    "bz_version": [],
    "bz_device_count": [C.POINTER(i32)],
    "bz_update_h_2d": [vp, i32, vp, i32, vp, i32, vp, i32, i32, i32, f64, f64, f64, f64, f64, vp],
    "bz_update_e_2d": [vp, i32, vp, i32, vp, i32, vp, i32, vp, i32, i32, i32, f64, f64, f64, f64, vp],
    "bz_update_3d": [P6, vp, vp, i32, i32, i32, i32, f64, f64, f64, f64, f64, f64, vp],
    "bz_yee3d_step": [P6, P6, vp, vp, i32, i32, i32, i32, f64, f64, f64, f64, f64, f64, C.POINTER(BoxSource), i32,
                      i32, i32, vp],
    "bz_yee3d_step2": [P6, P6, vp, vp, i32, i32, i32, i32, f64, f64, f64, f64, f64, f64, C.POINTER(BoxSource),
                       C.POINTER(f32), i32, vp, C.POINTER(i32), i32, i32, i32, vp],
    "bz_yee3d_step3": [P6, P6, vp, vp, i32, i32, i32, i32, f64, f64, f64, f64, f64, f64, C.POINTER(BoxSource),
                       C.POINTER(f32), C.POINTER(f32), i32, vp, vp, C.POINTER(i32), i32, i32, i32, vp],
    "bz_yee3d_step2_noalloc": [P6, P6, vp, vp, i32, i32, i32, i32, f64, f64, f64, f64, f64, f64,
                               C.POINTER(BoxSource), C.POINTER(f32), i32, C.POINTER(i32), vp, vp, vp, i32, i32, i32,
                               vp, vp, i32, vp, vp, vp, i32, i32, i32, vp],
    "bz_yee2d_run": [vp, i32, vp, i32, i32, vp, i32, vp, i32, vp, vp, i32, i32, vp, i32, vp, vp, vp],
    "bz_yee2d_hz_run": [vp, i32, vp, i32, i32, vp, i32, vp, i32, vp, vp, i32, i32, vp, i32, vp, vp, vp],
    "bz_yee2d_run2": [vp, i32, vp, i32, i32, vp, i32, i32, vp, vp, i32, vp, vp, i32, i32, vp, i32, vp, vp, vp],
    "bz_yee2d_coef": [vp, i32, i32, vp],
    "bz_gather2": [P6, vp, P6, vp, vp, i32, i32, i32, vp, vp, i32, vp, vp, vp, vp],
    "bz_gather3": [P6, P6, vp, P6, vp, vp, i32, i32, i32, i32, vp, vp, vp, i32, vp, vp, vp, vp, vp],
    "bz_abi_sizes": [C.POINTER(i32), i32],
    "bz_yee2d_runk": [vp, i32, vp, i32, i32, i32, i32, vp, i32, i32, vp, vp, vp, i32, vp, vp, i32, i32, vp, i32, vp,
                      vp, i32, vp],
    "bz_adj2d_run": [vp, i32, vp, i32, i32, i32, i32, vp],
    "bz_adj2d_run2": [vp, i32, vp, i32, i32, vp, i32, i32, i32, i32, i32, vp],
    "bz_adj2d_runk": [vp, i32, vp, i32, i32, i32, i32, vp, i32, i32, i32, i32, i32, vp],
    "bz_nn_conv3x3": [vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, vp],
    "bz_nn_conv3x3_dgrad": [vp, vp, vp, i32, i32, i32, i32, i32, vp],
    "bz_nn_conv3x3_dgrad_relu": [vp, vp, vp, vp, i32, i32, i32, i32, i32, vp],
    "bz_nn_wgrad_work": [i32, i32, i32, i32, i32, i32],
    "bz_nn_conv3x3_wgrad": [vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, vp, i64, vp],
    "bz_nn_first_work": [i32, i32, i32, i32],
    "bz_nn_conv_first": [vp, vp, vp, vp, i32, i32, i32, i32, vp],
    "bz_nn_conv_first_bwd": [vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, vp, i64, vp],
    "bz_nn_relu_mask": [vp, vp, i64, vp],
    "bz_nonfinite_count": [vp, i64, vp, vp],
    "bz_nn_avgpool2": [vp, vp, i32, i32, i32, i32, vp],
    "bz_nn_avgpool2_bwd": [vp, vp, i32, i32, i32, i32, vp],
    "bz_nn_avgpool2_bwd_relu": [vp, vp, vp, i32, i32, i32, i32, vp],
    "bz_nn_head": [vp, vp, vp, vp, vp, i32, i32, i32, i32, vp],
    "bz_nn_head_bwd": [vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, vp],
    "bz_nn_weights": [vp, vp, vp, i32, i32, vp],
    "bz_nn_weights_n": [vp, vp, vp, vp, vp, i32, vp],
    "bz_nn_adam": [vp, vp, vp, vp, i64, f32, f32, f32, f32, i32, vp],
    "bz_nn_to_bf16": [vp, vp, i64, vp],
    "bz_adj2d_inject": [vp, i32, i32, i32, vp],
    "bz_adj2d_grad": [vp, i32, i32, i32, i32, vp],
    "bz_adj2d_injtab": [vp, i32, i32, i32, i32, vp],
    "bz_filter2d": [vp, vp, i32, i32, i32, vp, i32, vp, vp, vp],
    "bz_add_points": [vp, vp, vp, i32, f32, vp],
    "bz_affine_points": [vp, vp, vp, vp, i32, f32, vp],
    "bz_gather": [P6, vp, vp, i32, vp, i32, vp, vp, vp],
    "bz_grid_power_2d": [vp, vp, vp, vp, i32, i32, i32, vp],
    "bz_grid_power_3d": [vp, P6, i32, i32, i32, i32, vp],
    "bz_grid_power_2d_cplx": [vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, vp],
    "bz_grid_power_3d_cplx": [vp, P6, P6, i32, i32, i32, i32, vp],
    "bz_raster_2d": [C.POINTER(Shape), i32, C.POINTER(f64), C.POINTER(Pml), i32, f64, f64, f64, i32, i32,
                     vp, vp, vp],
    "bz_raster_3d": [C.POINTER(Shape), i32, C.POINTER(f64), C.POINTER(Pml), i32, f64, f64, f64, f64, f64,
                     i32, i32, i32, vp, vp, vp],
}
_RESTYPE = {"bz_last_error": C.c_char_p, "bz_nn_wgrad_work": i64, "bz_nn_first_work": i64}

_lib = None


class NativeError(RuntimeError):
    pass


def lib():
    """Load (once) and return the library with argtypes set."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m beamz_amd.build`")
    # torch first: its import maps torch's own HIP runtime, which the library's HIP dependency then resolves
    # to. Loaded before torch (a mesh rasterised before any GPU work), the library would map the system HIP
    # runtime and torch a second one, and the library's launches fail with "no ROCm-capable device".
    import torch  # noqa: F401
    L = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, i32)
    _lib = L
    return L


def call(name, *args):
    """Call an entry point; raise ValueError/RuntimeError on a negative status."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().bz_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(msg)
        raise NativeError(msg or f"{name} failed with status {rc}")
    return rc


def ptrs6(seq):
    arr = P6()
    for q, p in enumerate(seq):
        arr[q] = p
    return arr
