"""ctypes binding of libskewhip.so (the C ABI in include/skewhip.h).

Loading fails loudly when the HIP library has not been built: there is no CPU fallback.
"""
import atexit
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# SKEWHIP_LIB names another build of the same library (the in-tree one by default)
LIB_PATH = os.environ.get("SKEWHIP_LIB") or os.path.join(_HERE, "libskewhip.so")

ENGINE_AUTO, ENGINE_LARGE, ENGINE_TEAM1, ENGINE_BLOCKED = 0, 1, 2, 3
FLAG_TIMING = 0x100
FLAG_PHASES = 0x200
ERR_NOCONV, ERR_DOMAIN, ERR_SINGULAR = 2, 6, 7


class SkhError(RuntimeError):
    """A non-zero status from libskewhip (message from skh_last_error)."""

    def __init__(self, func, code, msg):
        super().__init__("%s failed (status %d): %s" % (func, code, msg))
        self.code = code


class skh_opts(ctypes.Structure):
    _fields_ = [("nb", ctypes.c_int64), ("flags", ctypes.c_int64)]


_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_L = None


def lib():
    global _L
    if _L is not None:
        return _L
    if not os.path.exists(LIB_PATH):
        raise ImportError("libskewhip.so is not built (%s); run __graft_entry__.build() or "
                          "python skewlinearalgebra.jl_amd/build_lib.py" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    i64, vp, op = ctypes.c_int64, ctypes.c_void_p, ctypes.POINTER(skh_opts)
    sig = {
        "skh_init": ([ctypes.c_int], ctypes.c_int),
        "skh_finalize": ([], ctypes.c_int),
        "skh_version": ([], ctypes.c_char_p),
        "skh_last_error": ([], ctypes.c_char_p),
        "skh_sytrd_d": ([i64, vp, i64, vp, vp, _i64p, op], ctypes.c_int),
        "skh_sytrd_d_device": ([i64, vp, i64, vp, vp, vp, op], ctypes.c_int),
        "skh_sytrd_d_device_noalloc": ([i64, vp, i64, vp, vp, vp, op], ctypes.c_int),
        "skh_device_status": ([vp], ctypes.c_int),
        "skh_logpf_d": ([i64, vp, i64, _dp, _dp, op], ctypes.c_int),
        "skh_logpf_d_device": ([i64, vp, i64, vp, vp, vp, op], ctypes.c_int),
        "skh_logpf_batched_d": ([i64, vp, vp, vp, vp, vp, op], ctypes.c_int),
        "skh_logpf_batched_d_device": ([i64, vp, vp, vp, vp, vp, vp, op], ctypes.c_int),
        "skh_skewtrieigvals_d": ([i64, vp, vp], ctypes.c_int),
        "skh_skewtrieigen_d": ([i64, vp, vp, vp, vp, ctypes.c_int], ctypes.c_int),
        "skh_skewtrieigen_d_gpu": ([i64, vp, vp, vp, vp, ctypes.c_int], ctypes.c_int),
        "skh_skewfun_d": ([i64, vp, i64, ctypes.c_int, vp, vp, op], ctypes.c_int),
        "skh_skewtrieigvals_z": ([i64, vp, vp, vp], ctypes.c_int),
        "skh_skewtrieigen_z": ([i64, vp, vp, vp, vp], ctypes.c_int),
        "skh_skewfun_tri_d": ([i64, vp, ctypes.c_int, vp, vp], ctypes.c_int),
        "skh_skewfun_z": ([i64, vp, i64, ctypes.c_int, vp, vp, op], ctypes.c_int),
        "skh_diag_mfma_probe": ([vp], ctypes.c_int),
        "skh_diag_timings": ([vp], ctypes.c_int),
        "skh_diag_phase_clock": ([vp, ctypes.c_int], ctypes.c_int),
        "skh_orgtr_d_device": ([i64, vp, i64, vp, vp, i64, vp], ctypes.c_int),
        "skh_eigen_d": ([i64, vp, i64, vp, vp, vp, op], ctypes.c_int),
        "skh_logpf_grad_batched_d": ([i64, vp, vp, vp, vp, vp, vp, op], ctypes.c_int),
        "skh_logpf_grad_batched_d_device": ([i64, vp, vp, vp, vp, vp, vp, vp, op], ctypes.c_int),
        "skh_logpf_batched_plan_create": ([i64, vp, vp, op, ctypes.POINTER(vp)], ctypes.c_int),
        "skh_logpf_batched_plan_exec": ([vp, vp, vp, vp, vp], ctypes.c_int),
        "skh_logpf_batched_plan_tune": ([vp, vp, vp, vp, vp, _dp], ctypes.c_int),
        "skh_logpf_grad_batched_plan_create": ([i64, vp, vp, op, ctypes.POINTER(vp)], ctypes.c_int),
        "skh_logpf_grad_batched_plan_exec": ([vp, vp, vp, vp, vp, vp], ctypes.c_int),
        "skh_plan_status": ([vp, vp], ctypes.c_int),
        "skh_plan_info": ([vp, vp], ctypes.c_int),
        "skh_plan_config": ([vp, vp], ctypes.c_int),
        "skh_plan_set_config": ([vp, vp], ctypes.c_int),
        "skh_plan_destroy": ([vp], ctypes.c_int),
        "skh_sytrd_z": ([i64, vp, i64, vp, vp, vp, op], ctypes.c_int),
        "skh_sytrd_z_device": ([i64, vp, i64, vp, vp, vp, vp, op], ctypes.c_int),
        "skh_orgtr_z_device": ([i64, vp, i64, vp, vp, i64, vp], ctypes.c_int),
        "skh_eigvals_z": ([i64, vp, i64, vp, op], ctypes.c_int),
        "skh_eigen_z": ([i64, vp, i64, vp, vp, op], ctypes.c_int),
        "skh_skewchol_d": ([i64, vp, i64, vp, i64, vp, ctypes.c_int, vp, op], ctypes.c_int),
        "skh_skewchol_d_batched": ([i64, vp, vp, vp, vp, vp, vp, ctypes.c_int, op], ctypes.c_int),
        "skh_skewchol_d_batched_device": ([i64, vp, vp, vp, vp, vp, vp, ctypes.c_int, vp, op], ctypes.c_int),
        "skh_pfaffian_z_batched": ([i64, vp, vp, vp, vp, op], ctypes.c_int),
        "skh_pfaffian_z_batched_device": ([i64, vp, vp, vp, vp, vp, op], ctypes.c_int),
        "skh_logabspfaffian_z_batched": ([i64, vp, vp, vp, vp, vp, op], ctypes.c_int),
        "skh_logabspfaffian_z_batched_device": ([i64, vp, vp, vp, vp, vp, vp, op], ctypes.c_int),
        "skh_skewtrieigvals_range_d": ([i64, vp, i64, i64, ctypes.c_double, ctypes.c_double, ctypes.c_int, vp, _i64p],
                                       ctypes.c_int),
        "skh_eigvals_range_d": ([i64, vp, i64, i64, i64, ctypes.c_double, ctypes.c_double, ctypes.c_int, vp, _i64p, op],
                                ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _L = L
    # release the library's stream, BLAS handle and cached device buffers while the HIP runtime is
    # still up: left to the C exit handlers, they would be torn down after the runtime
    atexit.register(_finalize)
    return L


def _finalize():
    if _L is not None:
        _L.skh_finalize()


def check(func, code):
    if code != 0:
        msg = lib().skh_last_error().decode(errors="replace")
        if code < 0:
            raise ValueError("%s: invalid argument %d: %s" % (func, -code, msg))
        raise SkhError(func, code, msg)


def opts(nb=0, engine=ENGINE_AUTO):
    o = skh_opts()
    o.nb = int(nb)
    o.flags = int(engine)
    return o


def declared_symbols():
    """Function names declared in include/skewhip.h (used by the ABI export test)."""
    import re
    hdr = os.path.join(_HERE, "..", "..", "include", "skewhip.h")
    with open(hdr) as f:
        txt = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(skh_\w+)\s*\(", txt, re.M)))
