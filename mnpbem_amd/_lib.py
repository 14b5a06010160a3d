"""ctypes binding of libmnpbem_hip.so (declarations: include/mnpbem_hip.h).

The library is mandatory: importing this module raises when it is missing, and every call
raises ``RuntimeError`` on a non-zero status. There is no host fallback.
"""
import atexit
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmnpbem_hip.so")

P = C.c_void_p
I = C.c_int
L = C.c_long

# name -> argument types (all return int status)
_SIGS = {
    "mnp_abi_version": [],
    "mnp_handle_create": [P],
    "mnp_handle_destroy": [P],
    "mnp_handle_bind": [P],
    "mnp_release": [],
    "mnp_probe_mfma_layout": [P, P],
    "mnp_prof_enable": [I],
    "mnp_prof_collect": [P],
    "mnp_range_push": [C.c_char_p],
    "mnp_range_pop": [],
    "mnp_green_assemble": [I, I, I, I, I, P, P, P, P, P, P, P, I, P, P, P, P, I, P, P, P, P, L, L, P],
    "mnp_green_assemble_pair": [I, I, I, I, I, P, P, P, P, P, P, P, I, P, P, P, P, P, P, P, P, L, L, P, P, P, P, L, L,
                                P],
    "mnp_gemm_sub_batched": [I, I, I, I, P, L, L, P, L, L, P, L, L, I, P],
    "mnp_gemm_sub_batched_noalloc": [I, I, I, I, P, L, L, P, L, L, P, L, L, I, P],
    "mnp_noalloc_scratch_bytes": [P],
    "mnp_lu_reserve": [I, I, I, P],
    "mnp_getrf_batched": [I, I, I, P, L, L, P, L, P, I, P],
    "mnp_perm_batched": [I, I, P, L, P, P, L, P],
    "mnp_getrs_batched": [I, I, I, I, P, L, L, P, L, P, L, L, P, L, L, P],
    "mnp_rsolve_batched": [I, I, I, I, P, L, L, P, L, P, L, L, P, L, L, I, P],
    "mnp_gemv_batched": [I, I, I, I, P, L, L, P, L, L, P, L, L, P],
    "mnp_scale_diag_batched": [I, I, I, C.c_double, P, L, P, P, L, L, P],
    "mnp_bemret_factor": [I, I, I, P, P, P, P, P, P, P, L, L, P, P, P, P, L, P, P, P, I, P],
    "mnp_pad_batched": [I, I, I, P, P, L, L, C.c_double, C.c_double, P],
    "mnp_bemret_solve": [I, I, I, P, P, P, P, P, L, L, P, P, L, P, P, P, P, L, P, L, P, L, P],
    "mnp_bem_operator": [I, I, P, P, P, P, L, L, P, L, P, P, P, P, L, P, L, P, L, P, L, P, P],
    "mnp_planewave_rhs": [I, I, P, P, P, L, P, P, P, P, P, P, P, L, P],
    "mnp_farfield": [I, I, P, P, P, L, P, P, I, P, P, L, P, P],
    "mnp_mdot": [L, I, P, L, P, P, P],
    "mnp_mupdate": [L, I, P, L, P, P, P],
    "mnp_cast_batched": [I, L, I, P, L, P, L, P],
    "mnp_accum_batched": [I, L, I, P, L, P, L, P, P],
    "mnp_hmat_fill_dense": [P, P, I, P, P, P, L, P],
    "mnp_hmat_aca": [P, P, I, I, P, P, P, P, P, P, C.c_double, I, P],
    "mnp_hmat_compact": [I, P, I, P, P, P, P, P, P, P, P, P, P],
    "mnp_hmat_matvec": [P, I, I, P, L, P, L, P],
    "mnp_bem_operator_hmat": [P, P, P, P, P, P, P, P, P, L, P, P],
    "mnp_near_terms": [I, P, P, P, P, P, P, P, I, P, P, P],
    "mnp_colsum_far": [I, P, I, P, P, P, P, P, P],
    "mnp_mlp_gemm": [I, I, I, I, I, I, P, L, P, L, P, P, P, L, P],
    "mnp_mlp_gelu_bwd": [L, P, P, P, P],
    "mnp_mlp_colsum": [I, I, P, L, P, P],
    "mnp_mlp_mse": [L, P, P, P, P, P],
    "mnp_mlp_adam": [L, P, P, P, P, C.c_double, C.c_double, C.c_double, C.c_double, I, P],
    "mnp_probe_mfma_bf16": [P, P],
    "mnp_set_tuning": [I, I],
    "mnp_surface_eval": [P, I, C.c_double, P, P, P, P],
    "mnp_green_block": [P, I, I, I, I, P, P],
    "mnp_point_eval": [P, C.c_double, P, P, P, P, P, P],
    "mnp_green_pick": [P, I, I, P, C.c_long, P, P],
    "mnp_get_tuning": [I, P],
}


class GreenGeomC(C.Structure):
    """mnp_green_geom of include/mnpbem_hip.h (device pointers)."""
    _fields_ = [("n", I), ("npart", I), ("order", I)] + [
        (f, P) for f in ("pos", "nvec", "area", "fpart", "con_in", "con_out", "kmat", "rowptr", "rcol", "rg", "rf")]


class PointGeomC(C.Structure):
    """mnp_point_geom of include/mnpbem_hip.h (device pointers)."""
    _fields_ = [("npt", I), ("nface", I), ("npart", I), ("order", I)] + [
        (f, P) for f in ("ppos", "pmed", "fpos", "farea", "fpart", "inout", "kmat", "rowptr", "rcol", "rg", "rf")]


class HMatC(C.Structure):
    """mnp_hmat of include/mnpbem_hip.h (device pointers)."""
    _fields_ = [("n", I), ("nd", I), ("nlr", I), ("c2p", P), ("dblk", P), ("doff", P), ("D", P), ("dstride", L),
                ("lblk", P), ("uoff", P), ("voff", P), ("U", P), ("V", P), ("rank", P), ("tmp", P), ("toff", P)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("libmnpbem_hip.so not found at {}: build it with "
                          "`python -m mnpbem_amd._build` (hipcc, gfx950)".format(LIB_PATH))
    lib = C.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = I
    return lib


lib = _load()
# nothing of the library (workspaces, side streams, events) stays alive into the HIP runtime's
# and profilers' own shutdown
atexit.register(lib.mnp_release)


_STATUS = {2: "the operand planes of this shape exceed the static scratch (see mnp_noalloc_scratch_bytes)",
           3: "the static scratch's device address is unavailable",
           4: "workspaces cannot be reserved for this size (cooperative panels or allocation failure)"}


def call(name, *args):
    """Call ``name`` and raise on a non-zero status."""
    st = getattr(lib, name)(*args)
    if st != 0:
        if st < 0:
            raise ValueError("{}: argument {} rejected".format(name, -st))
        if st in _STATUS:
            raise ValueError("{}: {}".format(name, _STATUS[st]))
        raise RuntimeError("{}: HIP launch failure (status {})".format(name, st))
    return st


class stage:
    """roctx range around a host stage (``with stage("gmres_iter"):``), named "mnp:<name>" like
    the library's own per-entry-point ranges."""

    def __init__(self, name):
        self._name = ("mnp:" + name).encode()

    def __enter__(self):
        lib.mnp_range_push(self._name)
        return self

    def __exit__(self, *exc):
        lib.mnp_range_pop()


class Handle:
    """A library handle (knobs, workspaces, lookahead stream) for the calling thread:
    ``with Handle() as h:`` binds it for the block and restores the thread's default after;
    ``close()`` frees it."""

    def __init__(self):
        self._h = C.c_void_p()
        call("mnp_handle_create", C.byref(self._h))

    def bind(self):
        call("mnp_handle_bind", self._h)
        return self

    @staticmethod
    def unbind():
        call("mnp_handle_bind", None)

    def __enter__(self):
        return self.bind()

    def __exit__(self, *exc):
        self.unbind()

    def close(self):
        if self._h:
            call("mnp_handle_destroy", self._h)
            self._h = C.c_void_p()


def exported_symbols():
    return list(_SIGS)
