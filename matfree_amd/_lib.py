"""ctypes binding of the C ABI declared in ``include/slq.h``.

The shared library ``libslq.so`` is built in-tree by ``matfree_amd._build``.
There is no fallback: if the library is missing, every GPU entry point raises.
"""

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libslq.so")

# status codes (slq_status in include/slq.h)
SLQ_OK = 0
SLQ_EINVAL = -1
SLQ_EINVAL_K = -2
SLQ_ENOMEM = -3
SLQ_EHIP = -4
SLQ_EUNSUPPORTED = -5
SLQ_EADJOINT_K0 = -6
# bytes of the static device scratch of slq_logdet_noalloc (SLQ_NOALLOC_SCRATCH_BYTES)
NOALLOC_SCRATCH_BYTES = 256 << 20

# operator kinds / precisions
SLQ_OP_DENSE = 0
SLQ_OP_RBF = 1
PRECISIONS = {"fp32": 0, "bf16": 1, "bf16x3": 2, "bf16x2": 3}
# packed storage of a stored matrix (slq_storage)
STORAGES = {"fp32": 0, "bf16x2": 1, "bf16": 2, "fp8x2": 3, "fp8": 4}
# matrix functions of slq_funm_vectors (slq_matfun)
MATFUNS = {"log": 0, "exp": 1, "sin": 2, "cos": 3, "sqrt": 4, "reciprocal": 5}

c_void_p = ctypes.c_void_p
c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_u32 = ctypes.c_uint32


class OperatorStruct(ctypes.Structure):
    """Mirror of ``slq_operator``."""

    _fields_ = [
        ("kind", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("n", ctypes.c_int64),
        ("matrix", c_void_p),
        ("points", c_void_p),
        ("dim", ctypes.c_int32),
        ("storage", ctypes.c_int32),
        ("lengthscale", ctypes.c_double),
        ("noise", ctypes.c_double),
        ("packed", c_void_p),
        ("tile_scales", c_void_p),
    ]


# name -> (restype, argtypes); every function declared in include/slq.h
SIGNATURES = {
    "slq_context_create": (c_int, [c_int, c_void_p, ctypes.POINTER(c_void_p)]),
    "slq_context_destroy": (c_int, [c_void_p]),
    "slq_context_set_stream": (c_int, [c_void_p, c_void_p]),
    "slq_last_error": (ctypes.c_char_p, [c_void_p]),
    "slq_context_scratch_bytes": (c_i64, [c_void_p]),
    "slq_packed_size": (c_int, [c_i64, c_int, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64)]),
    "slq_pack_matrix": (c_int, [c_void_p, c_void_p, c_i64, c_int, c_void_p, c_void_p]),
    "slq_sampler_signs": (c_int, [c_void_p, c_u32, c_u32, c_i64, c_i64, c_void_p]),
    "slq_matvec": (c_int, [c_void_p, ctypes.POINTER(OperatorStruct), c_void_p, c_i64, c_void_p]),
    "slq_hessenberg": (
        c_int,
        [c_void_p, ctypes.POINTER(OperatorStruct), c_void_p, c_i64, c_i64, c_void_p, c_void_p,
         c_void_p, c_void_p],
    ),
    "slq_arnoldi": (
        c_int,
        [c_void_p, ctypes.POINTER(OperatorStruct), c_void_p, c_i64, c_i64, c_int, c_void_p,
         c_void_p, c_void_p, c_void_p],
    ),
    "slq_arnoldi_vjp": (
        c_int,
        [c_void_p, ctypes.POINTER(OperatorStruct), c_void_p, c_i64, c_i64, c_int, c_int, c_void_p,
         c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "slq_lanczos_vjp": (
        c_int,
        [c_void_p, ctypes.POINTER(OperatorStruct), c_void_p, c_i64, c_i64, c_void_p, c_void_p,
         c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "slq_logdet": (
        c_int,
        [c_void_p, ctypes.POINTER(OperatorStruct), c_void_p, c_i64, c_i64, c_void_p, c_void_p,
         c_void_p, c_void_p],
    ),
    "slq_logdet_noalloc": (
        c_int,
        [c_void_p, ctypes.POINTER(OperatorStruct), c_void_p, c_i64, c_i64, c_void_p, c_void_p,
         c_void_p, c_void_p],
    ),
    "slq_logdet_vjp": (
        c_int,
        [c_void_p, ctypes.POINTER(OperatorStruct), c_void_p, c_i64, c_i64, c_void_p, c_void_p,
         c_void_p, c_void_p],
    ),
    "slq_lanczos": (
        c_int,
        [c_void_p, ctypes.POINTER(OperatorStruct), c_void_p, c_i64, c_i64, c_void_p, c_void_p,
         c_void_p, c_void_p, c_void_p],
    ),
    "slq_logdet_lanczos": (
        c_int,
        [c_void_p, ctypes.POINTER(OperatorStruct), c_void_p, c_i64, c_i64, c_void_p, c_void_p,
         c_void_p, c_void_p],
    ),
    "slq_logdet_lanczos_vjp": (
        c_int,
        [c_void_p, ctypes.POINTER(OperatorStruct), c_void_p, c_i64, c_i64, c_void_p, c_void_p,
         c_void_p, c_void_p],
    ),
    "slq_funm_vectors": (
        c_int,
        [c_void_p, ctypes.POINTER(OperatorStruct), c_void_p, c_i64, c_i64, c_int, c_int,
         c_void_p],
    ),
    "slq_mean_and_sem": (c_int, [c_void_p, c_void_p, c_i64, c_void_p, c_void_p]),
    "slq_cholesky_partial": (
        c_int, [c_void_p, ctypes.POINTER(OperatorStruct), c_i64, c_int, c_void_p, c_void_p, c_void_p],
    ),
    "slq_lowrank_solve": (
        c_int, [c_void_p, c_void_p, c_i64, c_i64, ctypes.c_double, c_void_p, c_i64, c_void_p],
    ),
    "slq_cg": (
        c_int,
        [c_void_p, ctypes.POINTER(OperatorStruct), c_void_p, c_i64, ctypes.c_double,
         ctypes.c_double, c_i64, c_void_p, c_i64, ctypes.c_double, c_void_p, c_void_p],
    ),
    "slq_quadform_vjp": (
        c_int, [c_void_p, ctypes.POINTER(OperatorStruct), c_void_p, c_i64, c_void_p, c_void_p],
    ),
    "slq_context_release": (c_int, [c_void_p]),
    "slq_logdet_save": (
        c_int,
        [c_void_p, ctypes.POINTER(OperatorStruct), c_void_p, c_i64, c_i64, c_int, c_void_p,
         c_void_p, ctypes.POINTER(c_void_p)],
    ),
    "slq_logdet_vjp_saved": (
        c_int, [c_void_p, ctypes.POINTER(OperatorStruct), c_void_p, c_void_p, c_void_p],
    ),
    "slq_saved_bytes": (c_i64, [c_void_p]),
    "slq_saved_destroy": (c_int, [c_void_p, c_void_p]),
    "slq_workspace_bytes": (
        c_int, [c_void_p, c_int, ctypes.POINTER(OperatorStruct), c_i64, c_i64,
                ctypes.POINTER(c_i64)],
    ),
    "slq_workspace_growth": (
        c_int, [c_void_p, c_int, ctypes.POINTER(OperatorStruct), c_i64, c_i64,
                ctypes.POINTER(c_i64)],
    ),
    "slq_profile_enable": (c_int, [c_void_p, c_int]),
    "slq_profile_select": (c_int, [c_void_p, c_int]),
    "slq_profile_read": (c_int, [c_void_p, c_int, ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(c_i64)]),
}

# entry points sized by slq_workspace_bytes (slq_entry)
ENTRIES = {"logdet": 0, "logdet_lanczos": 1, "logdet_vjp": 2, "logdet_lanczos_vjp": 3}

# kernel classes of the measurement hooks (slq_kernel_class)
KERNEL_CLASSES = {"matvec": 0, "dots": 1, "project_dots": 2, "project": 3, "normalize": 4,
                  "quadrature": 5, "adjoint": 6}

_lib = None
_lock = threading.Lock()


def load():
    """Load libslq.so (raises ImportError with the build command if missing)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"{LIB_PATH} is not built; run `python -m matfree_amd._build` "
                    "(matfree_amd has no CPU fallback)"
                )
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


class SLQError(RuntimeError):
    """A HIP runtime failure inside libslq."""


def raise_for(rc, ctx_ptr):
    """Translate an slq_status into the reference's exception classes."""
    if rc == SLQ_OK:
        return
    msg = load().slq_last_error(ctx_ptr) if ctx_ptr else b""
    msg = (msg or b"").decode() or f"libslq status {rc}"
    if rc in (SLQ_EINVAL, SLQ_EINVAL_K, SLQ_EADJOINT_K0):
        raise ValueError(msg)
    if rc == SLQ_EUNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == SLQ_ENOMEM:
        raise MemoryError(msg)
    raise SLQError(msg)


class Context:
    """One ``slq_context`` per (device, thread)."""

    def __init__(self, device):
        lib = load()
        ptr = c_void_p()
        rc = lib.slq_context_create(int(device), None, ctypes.byref(ptr))
        if rc != SLQ_OK:
            raise SLQError(f"slq_context_create(device={device}) failed with status {rc}")
        self.ptr = ptr
        self.device = int(device)

    def set_stream(self, stream_handle):
        raise_for(load().slq_context_set_stream(self.ptr, c_void_p(stream_handle)), self.ptr)

    def call(self, name, *args):
        rc = getattr(load(), name)(self.ptr, *args)
        raise_for(rc, self.ptr)

    def profile(self, on, classes=None):
        """Event hooks on (all kernel classes, or only the named ones) or off."""
        if on and classes is not None:
            mask = 0
            for kc in classes:
                mask |= 1 << KERNEL_CLASSES[kc]
            raise_for(load().slq_profile_select(self.ptr, mask), self.ptr)
        else:
            raise_for(load().slq_profile_enable(self.ptr, 1 if on else 0), self.ptr)

    def profile_read(self, kernel_class):
        """(total device ms, launches) of one kernel class since the last read."""
        ms = ctypes.c_double()
        cnt = c_i64()
        rc = load().slq_profile_read(self.ptr, KERNEL_CLASSES[kernel_class], ctypes.byref(ms),
                                     ctypes.byref(cnt))
        raise_for(rc, self.ptr)
        return float(ms.value), int(cnt.value)

    def scratch_bytes(self):
        return int(load().slq_context_scratch_bytes(self.ptr))

    def release(self):
        """Free the context's pooled scratch (after its stream has drained)."""
        raise_for(load().slq_context_release(self.ptr), self.ptr)

    def workspace_bytes(self, entry, desc, P, k):
        """Scratch bytes one call of ``entry`` (see ENTRIES) needs for (desc, P, k)."""
        out = c_i64()
        rc = load().slq_workspace_bytes(self.ptr, ENTRIES[entry], ctypes.byref(desc), int(P),
                                        int(k), ctypes.byref(out))
        raise_for(rc, self.ptr)
        return int(out.value)

    def workspace_growth(self, entry, desc, P, k):
        """Bytes by which this context's pooled scratch grows in one call of
        ``entry`` for (desc, P, k), given what it holds now."""
        out = c_i64()
        rc = load().slq_workspace_growth(self.ptr, ENTRIES[entry], ctypes.byref(desc), int(P),
                                         int(k), ctypes.byref(out))
        raise_for(rc, self.ptr)
        return int(out.value)

    def close(self):
        if self.ptr:
            load().slq_context_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_contexts = threading.local()


def context(device, stream_handle):
    """The calling thread's context for ``device`` and ``stream_handle``.

    One context per (device, stream): a context's scratch is reused by every
    call made through it, in its stream's order, so work in flight on two
    streams at once (the serving loop's slots) needs a context each.
    """
    table = getattr(_contexts, "table", None)
    if table is None:
        table = _contexts.table = {}
    key = (int(device), int(stream_handle or 0))
    ctx = table.get(key)
    if ctx is None:
        ctx = table[key] = Context(device)
        ctx.set_stream(stream_handle)
    return ctx
