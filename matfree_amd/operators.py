"""Operators: the reference's ``matvec(v, *params)`` plugin, on the GPU.

matfree accepts any JAX-traceable ``matvec(v, *params)`` (decomp.py:460,
stochtrace.py:47-49).  This package runs the matvec inside its own HIP block
matvec, so a matvec must be one of the operator types below; a plain Python
callable is rejected with ``TypeError`` (there is no second code path).

* :data:`dense_matvec` -- ``matvec(v, A) = A @ v`` with the stored symmetric
  matrix as the parameter, as in tests/test_decomp/test_tridiag_sym.py:16-18.
* :class:`RBF` -- ``matvec(v, lengthscale, noise) = K(x; l) v + noise * v``
  with ``K_ij = exp(-|x_i - x_j|^2 / (2 l^2))``, the covariance of
  tutorials/8_gaussian_logpdf.py:21-31, generated tile by tile on the fly.

* :data:`packed_matvec` -- ``matvec(v, M) = M @ v`` for a stored matrix
  packed once by :func:`pack` into a low-precision form (BF16 hi/lo planes,
  BF16, FP8 E4M3 with per-tile scales, or two FP8 planes): the stored-matrix
  variant whose block matvec reads 1-4 bytes per entry instead of 4.

All are callables: ``op(v, *params)`` applies the block matvec to a vector
``(n,)`` or a batch ``(P, n)`` and returns a float32 tensor on the GPU.
"""

import ctypes

import torch

from . import _lib, _runtime


class Operator:
    """Base class: builds the ``slq_operator`` descriptor for given parameters."""

    kind = None
    num_params = 0

    def __init__(self, precision="fp32"):
        if precision not in _lib.PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_lib.PRECISIONS)}, got {precision!r}")
        self.precision = precision

    def descriptor(self, n, params, device):
        """Return ``(OperatorStruct, keepalive)`` for dimension ``n``."""
        raise NotImplementedError

    def __call__(self, v, *params):
        device = _runtime.target_device(v, *params)
        x = _runtime.as_f32(v, device)
        single = x.ndim == 1
        X = x.reshape(1, -1) if single else x
        P, n = X.shape
        desc, keep = self.descriptor(n, params, device)
        W = torch.empty_like(X)
        ctx = _runtime.context(device)
        ctx.call("slq_matvec", ctypes.byref(desc), _runtime.ptr(X), P, _runtime.ptr(W))
        del keep
        return W[0] if single else W


class DenseMatvec(Operator):
    """``matvec(v, A) = A @ v`` for a stored symmetric ``A`` (the parameter)."""

    kind = _lib.SLQ_OP_DENSE
    num_params = 1

    def descriptor(self, n, params, device):
        if len(params) != 1:
            raise TypeError(f"dense_matvec expects one parameter (the matrix), got {len(params)}")
        A = _runtime.as_f32(params[0], device)
        if A.shape != (n, n):
            raise ValueError(f"matrix of shape {tuple(A.shape)} does not act on vectors of size {n}")
        d = _lib.OperatorStruct()
        d.kind = self.kind
        d.precision = _lib.PRECISIONS[self.precision]
        d.n = n
        d.matrix = A.data_ptr()
        d.points = None
        d.dim = 0
        return d, A


dense_matvec = DenseMatvec()


class RBF(Operator):
    """Matrix-free squared-exponential kernel plus nugget on points ``x`` (n, dim)."""

    kind = _lib.SLQ_OP_RBF
    num_params = 2

    def __init__(self, points, precision="fp32"):
        super().__init__(precision)
        pts = points if isinstance(points, torch.Tensor) else torch.as_tensor(points)
        if pts.ndim == 1:
            pts = pts[:, None]
        if pts.ndim != 2 or not (1 <= pts.shape[1] <= 4):
            raise ValueError("RBF points must have shape (n, dim) with 1 <= dim <= 4")
        self.points = pts
        self.n = int(pts.shape[0])
        self.dim = int(pts.shape[1])
        self._cache = {}

    def _points_on(self, device):
        key = str(device)
        t = self._cache.get(key)
        if t is None:
            t = self._cache[key] = _runtime.as_f32(self.points, device)
        return t

    def descriptor(self, n, params, device):
        if len(params) != 2:
            raise TypeError(f"RBF matvec expects (lengthscale, noise), got {len(params)} parameters")
        if n != self.n:
            raise ValueError(f"RBF operator on {self.n} points does not act on vectors of size {n}")
        X = self._points_on(device)
        d = _lib.OperatorStruct()
        d.kind = self.kind
        d.precision = _lib.PRECISIONS[self.precision]
        d.n = n
        d.matrix = None
        d.points = X.data_ptr()
        d.dim = self.dim
        d.lengthscale = _runtime.scalar(params[0])
        d.noise = _runtime.scalar(params[1])
        return d, X


class PackedMatrix:
    """A stored symmetric ``n x n`` matrix in a packed form of ``slq_storage``.

    ``storage``: ``"bf16x2"`` (A = hi + lo, two BF16 planes, ~16 significant
    bits), ``"bf16"`` (hi only), ``"fp8x2"`` (two OCP FP8 E4M3 planes with one
    power-of-two scale per 32 x 32 tile and plane, ~8 bits) or ``"fp8"`` (one
    plane, 4 bits).  ``packed`` (uint8) and ``scales`` (float32, FP8 forms)
    are device tensors in the layout documented in include/slq.h."""

    def __init__(self, A, storage="bf16x2"):
        if storage not in _lib.STORAGES or storage == "fp32":
            raise ValueError(f"storage must be one of "
                             f"{sorted(k for k in _lib.STORAGES if k != 'fp32')}, got {storage!r}")
        device = _runtime.target_device(A)
        A32 = _runtime.as_f32(A, device)
        if A32.ndim != 2 or A32.shape[0] != A32.shape[1]:
            raise ValueError(f"expected a square matrix, got shape {tuple(A32.shape)}")
        n = int(A32.shape[0])
        code = _lib.STORAGES[storage]
        nbytes, nscales = ctypes.c_int64(), ctypes.c_int64()
        rc = _lib.load().slq_packed_size(n, code, ctypes.byref(nbytes), ctypes.byref(nscales))
        if rc != _lib.SLQ_OK:
            raise ValueError(f"cannot pack an {n} x {n} matrix as {storage}")
        self.n = n
        self.storage = storage
        self.device = device
        self.packed = torch.empty((nbytes.value,), dtype=torch.uint8, device=device)
        self.scales = (torch.empty((nscales.value,), dtype=torch.float32, device=device)
                       if nscales.value else None)
        ctx = _runtime.context(device)
        ctx.call("slq_pack_matrix", _runtime.ptr(A32), n, code, _runtime.ptr(self.packed),
                 _runtime.ptr(self.scales))

    def nbytes(self):
        return self.packed.numel() + (4 * self.scales.numel() if self.scales is not None else 0)


def pack(A, storage="bf16x2"):
    """Pack a stored symmetric matrix for :data:`packed_matvec`."""
    return PackedMatrix(A, storage)


class PackedDense(Operator):
    """``matvec(v, M) = M @ v`` for a :class:`PackedMatrix` ``M`` (the
    parameter): BF16 MFMAs on the stored planes, vectors rounded to BF16,
    FP32 accumulation."""

    kind = _lib.SLQ_OP_DENSE
    num_params = 1

    def descriptor(self, n, params, device):
        if len(params) != 1:
            raise TypeError(f"packed_matvec expects one parameter (the packed matrix), "
                            f"got {len(params)}")
        M = params[0]
        if not isinstance(M, PackedMatrix):
            raise TypeError(f"packed_matvec expects a PackedMatrix (operators.pack), "
                            f"got {type(M).__name__}")
        if M.n != n:
            raise ValueError(f"matrix of shape ({M.n}, {M.n}) does not act on vectors of size {n}")
        if M.device != device:
            raise ValueError(f"packed matrix on {M.device}, vectors on {device}")
        d = _lib.OperatorStruct()
        d.kind = self.kind
        d.precision = _lib.PRECISIONS["fp32"]
        d.n = n
        d.matrix = None
        d.points = None
        d.dim = 0
        d.storage = _lib.STORAGES[M.storage]
        d.packed = M.packed.data_ptr()
        d.tile_scales = M.scales.data_ptr() if M.scales is not None else None
        return d, M


packed_matvec = PackedDense()


def resolve(matvec):
    """Check that ``matvec`` is a GPU operator of this package."""
    if not isinstance(matvec, Operator):
        raise TypeError(
            "matfree_amd runs the matvec inside its HIP block matvec; pass one of its "
            f"operators (dense_matvec, packed_matvec, RBF(points)), not {type(matvec).__name__}"
        )
    return matvec
