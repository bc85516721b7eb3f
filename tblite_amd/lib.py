"""ctypes binding of ``libtbx_hip.so`` (the C-ABI of ``include/tbx.h``).

The library is built in-tree by ``make -C tblite_amd`` (``__graft_entry__.build()``).
There is no CPU fallback: every entry point here runs the gfx950 kernels, and
loading fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as ct
import re
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libtbx_hip.so"
HEADER = HERE.parent / "include" / "tbx.h"

TBX_MAXG = 12
TBX_MAXSH = 4


class TbxError(RuntimeError):
    """Non-zero status of a C-ABI call; ``code`` is the TBX_ERR_* value."""

    def __init__(self, code, msg):
        super().__init__(f"tbx error {code}: {msg}")
        self.code = code


class Basis(ct.Structure):
    _fields_ = [("nid", ct.c_int), ("msh", ct.c_int), ("nsh_id", ct.c_void_p),
                ("ang", ct.c_void_p), ("nprim", ct.c_void_p), ("alpha", ct.c_void_p),
                ("coeff", ct.c_void_p), ("intcut", ct.c_double)]


class H0Par(ct.Structure):
    _fields_ = [("hscale", ct.c_void_p), ("shpoly", ct.c_void_p), ("rad", ct.c_void_p),
                ("ksig", ct.c_void_p), ("kpi", ct.c_void_p), ("kdel", ct.c_void_p)]


class Geom(ct.Structure):
    _fields_ = [("nat", ct.c_int), ("id", ct.c_void_p), ("xyz", ct.c_void_p),
                ("ntrans", ct.c_int), ("trans", ct.c_void_p), ("cutoff", ct.c_double)]


class List(ct.Structure):
    _fields_ = [("inl", ct.c_void_p), ("nnl", ct.c_void_p), ("nnbr", ct.c_int),
                ("nlat", ct.c_void_p), ("nltr", ct.c_void_p)]


class BatchDesc(ct.Structure):
    _fields_ = [("nmol", ct.c_int), ("mol_nat", ct.c_void_p), ("id", ct.c_void_p),
                ("bas", ct.POINTER(Basis)), ("h0", ct.POINTER(H0Par)), ("cutoff", ct.c_double)]


class BatchSizes(ct.Structure):
    _fields_ = [("nmol", ct.c_int), ("nat", ct.c_int), ("nsh", ct.c_int), ("nao", ct.c_int),
                ("ndense", ct.c_longlong), ("npair", ct.c_longlong)]


_lib = None

_P = ct.c_void_p
class Tiles(ct.Structure):
    _fields_ = [("fwd_unit_tasks", ct.c_int), ("fwd_unit_pairs", ct.c_int), ("fwd_chunk", ct.c_int),
                ("grad_unit_tasks", ct.c_int), ("grad_unit_pairs", ct.c_int),
                ("grad_chunk", ct.c_int), ("fwd_band", ct.c_int), ("grad_band", ct.c_int)]


class PlanStats(ct.Structure):
    _fields_ = [(n, ct.c_longlong) for n in (
        "nunit_fwd", "nunit_grad", "npair_fwd", "npair_grad", "ntask_fwd", "ntask_grad",
        "nelem", "ndense", "plan_bytes_fwd", "plan_bytes_grad", "bad_pairs", "bad_elems",
        "bad_units", "nunit_grad_live")]


_SIGS = {
    "tbx_version": (ct.c_int, []),
    "tbx_last_error": (ct.c_char_p, []),
    "tbx_context_new": (_P, [ct.c_int]),
    "tbx_context_delete": (None, [_P]),
    "tbx_context_set_memory_limit": (ct.c_int, [_P, ct.c_longlong]),
    "tbx_context_memory_used": (ct.c_longlong, [_P]),
    "tbx_get_selfenergy": (ct.c_int, [ct.c_int, _P, _P, _P, ct.c_int, _P, _P, _P, _P, _P]),
    "tbx_get_lattice_points": (ct.c_int, [_P, _P, ct.c_double, ct.POINTER(ct.c_int), _P]),
    "tbx_batch_create_periodic": (ct.c_int, [_P, ct.POINTER(BatchDesc), _P, _P, ct.POINTER(_P)]),
    "tbx_get_hamiltonian": (ct.c_int, [_P, ct.POINTER(Geom), ct.POINTER(List), ct.POINTER(Basis),
                                       ct.POINTER(H0Par), _P, _P, ct.c_int, _P, _P, _P, _P, _P]),
    "tbx_get_multipole_integrals": (ct.c_int, [_P, ct.POINTER(Geom), ct.POINTER(Basis), _P, _P,
                                               ct.c_int, _P, _P, _P]),
    "tbx_get_dipole_integrals": (ct.c_int, [_P, ct.POINTER(Geom), ct.POINTER(Basis), _P, _P,
                                            ct.c_int, _P, _P]),
    "tbx_get_overlap": (ct.c_int, [_P, ct.POINTER(Geom), ct.POINTER(Basis), _P, _P, ct.c_int,
                                   _P]),
    "tbx_get_hamiltonian_gradient": (ct.c_int, [_P, ct.POINTER(Geom), ct.POINTER(List),
                                                ct.POINTER(Basis), ct.POINTER(H0Par), _P, _P,
                                                ct.c_int, _P, _P, ct.c_int, _P, _P, _P, _P, _P,
                                                _P, _P, _P]),
    "tbx_get_hamiltonian_adjoint": (ct.c_int, [_P, ct.POINTER(Geom), ct.POINTER(List),
                                               ct.POINTER(Basis), ct.POINTER(H0Par), _P, _P,
                                               ct.c_int, _P, _P, _P, _P, _P, _P, _P, _P]),
    "tbx_batch_add_vao_to_h1": (ct.c_int, [_P, ct.c_int, _P, _P, _P, _P]),
    "tbx_batch_add_vmp_to_h1": (ct.c_int, [_P, ct.c_int, ct.c_int, _P, ct.c_int, _P, _P, _P]),
    "tbx_batch_mulliken_shell_charges": (ct.c_int, [_P, ct.c_int, _P, _P, _P, _P, _P]),
    "tbx_batch_mulliken_atomic_multipoles": (ct.c_int, [_P, ct.c_int, ct.c_int, _P, _P, _P,
                                                        _P]),
    "tbx_batch_last_scf_timing": (ct.c_int, [_P, _P]),
    "tbx_batch_create": (ct.c_int, [_P, ct.POINTER(BatchDesc), ct.POINTER(_P)]),
    "tbx_batch_delete": (None, [_P]),
    "tbx_batch_sizes_get": (ct.c_int, [_P, ct.POINTER(BatchSizes)]),
    "tbx_batch_offsets": (ct.c_int, [_P, _P, _P, _P, _P]),
    "tbx_batch_forward": (ct.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "tbx_batch_gradient": (ct.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "tbx_batch_gradient_with_forward": (ct.c_int, [_P] * 15),
    "tbx_batch_gradient_noalloc": (ct.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                              _P]),
    "tbx_batch_adjoint": (ct.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "tbx_batch_adjoint_with_forward": (ct.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                                  _P]),
    "tbx_batch_last_timing": (ct.c_int, [_P, ct.POINTER(ct.c_double), ct.POINTER(ct.c_double),
                                         ct.POINTER(ct.c_double), ct.POINTER(ct.c_double)]),
    "tbx_probe_copy": (ct.c_int, [_P, _P, ct.c_longlong, _P]),
    "tbx_batch_set_tiles": (ct.c_int, [_P, ct.POINTER(Tiles)]),
    "tbx_batch_get_tiles": (ct.c_int, [_P, ct.POINTER(Tiles)]),
    "tbx_batch_plan_stats": (ct.c_int, [ct.POINTER(BatchDesc), ct.POINTER(Tiles), _P,
                                        ct.c_double, ct.POINTER(PlanStats)]),
    "tbx_batch_order_pairs": (ct.c_int, [_P, _P, ct.c_double]),
    "tbx_batch_sygvd": (ct.c_int, [_P, _P, _P, _P, _P]),
    "tbx_batch_fermi_filling": (ct.c_int, [_P, _P, ct.c_double, _P, _P, _P, _P]),
    "tbx_batch_get_density_matrix": (ct.c_int, [_P, _P, _P, _P, _P]),
    "tbx_batch_get_wdensity": (ct.c_int, [_P, _P, _P, _P, _P, _P]),
    "tbx_batch_set_ncoord": (ct.c_int, [_P, _P, ct.c_double, ct.c_double, ct.c_double,
                                        ct.c_double]),
    "tbx_batch_ncoord": (ct.c_int, [_P, _P, _P, _P]),
    "tbx_batch_ncoord_gradient": (ct.c_int, [_P, _P, _P, _P, _P, _P]),
}


def header_functions():
    """Names of the functions ``include/tbx.h`` declares."""
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(tbx_\w+)\s*\(", text)))


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} not built: run `make -C {HERE}` "
                               "(no CPU fallback exists for this path)")
        L = ct.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != 0:
        raise TbxError(status, lib().tbx_last_error().decode())


_ctx = {}


def context(device: int = 0):
    """Per-device context, created on first use."""
    if device not in _ctx:
        c = lib().tbx_context_new(device)
        if not c:
            raise TbxError(3, lib().tbx_last_error().decode())
        _ctx[device] = c
    return _ctx[device]
