"""ctypes binding of libfdtd_hip.so (declarations in include/fdtd.h).

The library is the only compute path of this package: if it is missing or
fails to load, every call raises; nothing falls back to a CPU or PyTorch
implementation.
"""
from __future__ import annotations

import ctypes as C
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfdtd_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "fdtd.h")

FDTD_OK = 0
FDTD_EINVAL = -1
MAX_SOURCES = 16
MAX_PROBES = 16
MAX_CYLINDERS = 64
MAX_FREQ = 16
ENERGY_WORK = 1025
SRC_NONE, SRC_HARD, SRC_ADD = 0, 1, 2


class FdtdError(RuntimeError):
    """A call into libfdtd_hip.so failed (message from fdtd_last_error)."""

    def __init__(self, func: str, code: int, msg: str):
        super().__init__(f"{func} failed ({code}): {msg}")
        self.code = code


_f = C.POINTER(C.c_float)


class FdtdPml(C.Structure):
    """fdtd_pml: the pmlayer profile set (fd2d/pycuda/fd2d_3_2.py:204-209)."""
    _fields_ = [(k, _f) for k in ("fx1", "fx2", "fx3", "fy1", "fy2", "fy3",
                                   "gx2", "gx3", "gy2", "gy3")]


class FdtdMedium(C.Structure):
    """fdtd_medium: naz, nbz (fd2d/pycuda/fd2d_3_4.py:96-98)."""
    _fields_ = [("naz", _f), ("nbz", _f)]


class Fd2dTmDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int),
        ("dz", _f), ("iz", _f), ("ez", _f), ("hx", _f), ("hy", _f), ("ihx", _f), ("ihy", _f),
        ("dz2", _f), ("iz2", _f), ("hx2", _f), ("hy2", _f), ("ihx2", _f), ("ihy2", _f),
        ("md", FdtdMedium),
        ("pml", FdtdPml),
        ("nsrc", C.c_int),
        ("src_i", C.c_int * MAX_SOURCES), ("src_j", C.c_int * MAX_SOURCES),
        ("src_wave", _f), ("wave_len", C.c_int),
        ("nprobe", C.c_int),
        ("probe_i", C.c_int * MAX_PROBES), ("probe_j", C.c_int * MAX_PROBES),
        ("probe_out", _f), ("probe_len", C.c_int),
        ("depth", C.c_int), ("rows", C.c_int), ("cols", C.c_int),
        ("band_rows", C.c_int * 2), ("band_cols", C.c_int * 2),
        ("form", C.c_int),
        ("tile_work", C.c_void_p),
        ("tile_elems", C.c_size_t),
        ("tile_keep", C.c_int),
    ]


TM_FORM_AUTO, TM_FORM_STRIP, TM_FORM_MULTIWAVE = 0, 1, 2


class Fd2dGrid(C.Structure):
    """fd2d_grid: one grid of a batch (include/fdtd.h)."""
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int),
        ("dz", _f), ("iz", _f), ("ez", _f), ("hx", _f), ("hy", _f), ("ihx", _f), ("ihy", _f),
        ("dz2", _f), ("iz2", _f), ("hx2", _f), ("hy2", _f), ("ihx2", _f), ("ihy2", _f),
        ("md", FdtdMedium),
        ("pml", FdtdPml),
        ("band_rows", C.c_int * 2), ("band_cols", C.c_int * 2),
        ("src_i", C.c_int), ("src_j", C.c_int),
        ("src_wave", _f), ("wave_len", C.c_int),
        ("probe_i", C.c_int), ("probe_j", C.c_int),
        ("probe_out", _f), ("probe_len", C.c_int),
    ]


class FdtdFtrans(C.Structure):
    """fdtd_ftrans: the DFT planes (fd2d/pycuda/fd2d_3_4.py:100-103)."""
    _fields_ = [(k, _f) for k in ("r_pt", "i_pt", "r_in", "i_in")]


class FdtdMedium1d(C.Structure):
    """fdtd_medium1d: nax, nbx, ncx, ndx (fd1d/pycuda/fd1d_2_3.py:68-71)."""
    _fields_ = [(k, _f) for k in ("nax", "nbx", "ncx", "ndx")]


class Fd1dDformDesc(C.Structure):
    """fd1d_dform_desc: a 1-D D/E-form run (include/fdtd.h)."""
    _fields_ = [
        ("nx", C.c_int),
        ("dx", _f), ("ix", _f), ("sx", _f), ("ex", _f), ("hy", _f), ("bc", _f),
        ("md", FdtdMedium1d),
        ("src_pos", C.c_int), ("src_wave", _f), ("wave_len", C.c_int),
        ("nf", C.c_int), ("freq", _f), ("dt", C.c_float),
        ("ft", FdtdFtrans),
        ("nprobe", C.c_int), ("probe_k", C.c_int * MAX_PROBES),
        ("probe_out", _f), ("probe_len", C.c_int),
        ("work", _f), ("work_elems", C.c_size_t),
    ]


class Fd2dTfsfDesc(C.Structure):
    """fd2d_tfsf_desc: incident plane wave and DFT of a fd2d_tfsf_run."""
    _fields_ = [
        ("npml", C.c_int),
        ("ezi", _f), ("hxi", _f), ("bc", _f),
        ("inc_pos", C.c_int), ("inc_wave", _f), ("inc_len", C.c_int),
        ("nf", C.c_int), ("freq", _f), ("dt", C.c_float),
        ("ft", FdtdFtrans), ("dft_probe", C.c_int),
        ("work", _f), ("work_elems", C.c_size_t),
    ]


class Fd2dLbcDesc(C.Structure):
    """fd2d_lbc_desc: network and boundary-sample history (include/fdtd.h)."""
    _fields_ = [("params", C.c_void_p), ("wbuf", C.c_void_p), ("hist", C.c_void_p),
                ("hist2", C.c_void_p), ("work", C.c_void_p)]


class Fd2dBatchDesc(C.Structure):
    _fields_ = [
        ("ngrids", C.c_int),
        ("grids_dev", C.c_void_p), ("units_dev", C.c_void_p),
        ("nunits", C.c_int), ("depth", C.c_int), ("rows", C.c_int), ("cols", C.c_int),
        ("waves", C.c_int), ("vec", C.c_int),
        ("wave_len", C.c_int), ("probe_len", C.c_int),
        ("unit_work", C.c_void_p),
    ]


_P = C.c_void_p  # device pointers and streams are passed as integers
_I = C.c_int
_D = C.c_double
_F = C.c_float
_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)

# name -> (restype, argtypes)
SIGNATURES = {
    "fdtd_abi_version": (_I, []),
    "fdtd_last_error": (C.c_char_p, []),
    "fdtd_preload": (_I, []),
    "fdtd_gaussian_f32": (_F, [_I, _I, _F]),
    "fdtd_gaussian_f64": (_D, [_I, _I, _D]),
    "fdtd_sinusoidal_f32": (_F, [_I, _F, _F]),
    "fdtd_sinusoidal_f64": (_D, [_I, _D, _D]),
    "fd1d_exfield": (_I, [_I, _I, _P, _P, _P, _P, _I, _I, _F, _P]),
    "fd1d_exfield_f64": (_I, [_I, _I, _P, _P, _P, _P, _I, _I, _D, _P]),
    "fd1d_hyfield": (_I, [_I, _P, _P, _P, _P]),
    "fd1d_hyfield_f64": (_I, [_I, _P, _P, _P, _P]),
    "fd1d_run": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _I, _I, _P, _I, _IP, _P, _P]),
    "fd1d_run_f64": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _I, _I, _P, _I, _IP, _P, _P]),
    "fd1d_run_max_nx": (_I, [_I]),
    "fd1d_dielectric": (_I, [_I, _D, _D, _D, _P, _P, _P]),
    "fd1d_dielectric_f64": (_I, [_I, _D, _D, _D, _P, _P, _P]),
    "fd2d_dfield": (_I, [_I, _I, _I, C.POINTER(FdtdPml), _P, _P, _P, _I, _IP, _IP,
                         C.POINTER(C.c_float), _P]),
    "fd2d_efield": (_I, [_I, _I, C.POINTER(FdtdMedium), _P, _P, _P, _P]),
    "fd2d_hfield": (_I, [_I, _I, C.POINTER(FdtdPml), _P, _P, _P, _P, _P, _P]),
    "fd2d_pmlparam": (_I, [_I, _I, _I, C.POINTER(FdtdPml), _P]),
    "fd2d_dielectric": (_I, [_I, _I, _I, _D, _I, _DP, _I, _P, _P, _P]),
    "fd2d_energy": (_I, [_I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _DP, _P]),
    "fd2d_energy_noalloc": (_I, [_I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P]),
    "fd2d_energy_diff": (_I, [_I, _I, _I, _I, _I, _I, _P, _P, _P, _I, _I, _I, _I, _P, _P, _P, _P,
                              _DP, _P]),
    "fd2d_tm_run": (_I, [C.POINTER(Fd2dTmDesc), _I, _I, _P]),
    "fd2d_tm_plan": (_I, [C.POINTER(Fd2dTmDesc), _I, _IP, _IP, _IP, _IP]),
    "fd2d_tm_tile_elems": (C.c_size_t, [C.POINTER(Fd2dTmDesc), _I]),
    "fd2d_batch_units": (_I, [C.POINTER(Fd2dGrid), _I, _I, _I, _I, _IP]),
    "fd2d_batch_upload": (_I, [C.POINTER(Fd2dGrid), _I, _I, _I, _I, _I, _P, _P, _I,
                               C.POINTER(Fd2dBatchDesc), _P]),
    "fd2d_batch_run": (_I, [C.POINTER(Fd2dBatchDesc), _I, _I, _I, _I, _P]),
    "fd2d_batch_graph": (_I, [C.POINTER(Fd2dBatchDesc), _I, _I, _I, _I, _P, C.POINTER(C.c_void_p)]),
    "fdtd_graph_launch": (_I, [_P, _P]),
    "fdtd_graph_destroy": (_I, [_P]),
    "fd1d_dxfield": (_I, [_I, _I, _P, _P, _I, _F, _P]),
    "fd1d_exfield_dform": (_I, [_I, C.POINTER(FdtdMedium1d), _P, _P, _P, _P, _P]),
    "fd1d_fourier": (_I, [_I, _I, _I, _F, C.POINTER(C.c_float), _P, C.POINTER(FdtdFtrans), _P]),
    "fd1d_dielectric_dform": (_I, [_I, _F, _F, _F, _F, _F, C.POINTER(FdtdMedium1d), _P]),
    "fd1d_dform_run": (_I, [C.POINTER(Fd1dDformDesc), _I, _I, _P]),
    "fd1d_dform_work_elems": (C.c_size_t, [_I]),
    "fd2d_ezinct": (_I, [_I, _P, _P, _P, _P]),
    "fd2d_dfield_inc": (_I, [_I, _I, _I, C.POINTER(FdtdPml), _P, _P, _P, _P, _I, _F, _P]),
    "fd2d_inctdz": (_I, [_I, _I, _I, _P, _P, _P]),
    "fd2d_hxinct": (_I, [_I, _P, _P, _P]),
    "fd2d_incthx": (_I, [_I, _I, _I, _P, _P, _P]),
    "fd2d_incthy": (_I, [_I, _I, _I, _P, _P, _P]),
    "fd2d_fourier": (_I, [_I, _I, _I, _I, _F, C.POINTER(C.c_float), _P, _P, C.POINTER(FdtdFtrans), _P]),
    "fd2d_tfsf_run": (_I, [C.POINTER(Fd2dTmDesc), C.POINTER(Fd2dTfsfDesc), _I, _I, _P]),
    "fd2d_tfsf_work_elems": (C.c_size_t, [_I, _I]),
    "lbc_hist_elems": (C.c_size_t, [_I, _I]),
    "lbc_work_bytes": (C.c_size_t, [_I, _I]),
    "fd2d_lbc_margin": (_I, []),
    "fd2d_lbc_run": (_I, [C.POINTER(Fd2dTmDesc), C.POINTER(Fd2dLbcDesc), _I, _I, _P]),
    "lbc_param_count": (_I, []),
    "lbc_wbuf_elems": (_I, []),
    "lbc_train_smem": (C.c_size_t, []),
    "lbc_sample": (_I, [_I, _I, _I, _P, _P, _P, _P, _P]),
    "lbc_pack": (_I, [_P, _P, _P]),
    "lbc_forward": (_I, [_I, _P, _P, C.c_longlong, _P, _P, _P, _P, _P]),
    "lbc_train_grad": (_I, [_I, _P, _P, C.c_longlong, _P, _F, _P, _P, _I, _P, _P, _F, C.c_uint,
                            _P]),
    "lbc_adam": (_I, [_I, _P, _P, _P, _P, _I, _F, _F, _F, _F, _P, _P, _I, _I, _P]),
    "lbc_gram_elems": (_I, []),
    "lbc_gram": (_I, [_I, _P, _P, C.c_longlong, _P, _I, _F, _I, _P, _P]),
    "lbc_fit_linear": (_I, [_I, _P, _P, C.c_double, _P, _P, _P]),
}

_lib = None


def header_symbols() -> list[str]:
    """Names of the functions declared in include/fdtd.h."""
    txt = open(HEADER_PATH).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", txt, flags=re.M)))


def lib():
    """Load libfdtd_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m simulation_amd.build` "
                "(the HIP library is the only compute path)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


_preloaded = False


def preload() -> None:
    """Load every device code object of the library (once per process), so
    that no later launch carries a code-object load (fdtd_preload)."""
    global _preloaded
    if not _preloaded:
        call("fdtd_preload")
        _preloaded = True


def check(name: str, rc: int) -> None:
    if rc != FDTD_OK:
        msg = lib().fdtd_last_error().decode(errors="replace")
        raise FdtdError(name, rc, msg)


def call(name: str, *args) -> None:
    check(name, getattr(lib(), name)(*args))
