"""The C ABI: libslq.so loads, exports every function of include/slq.h, and
the ctypes mirror of ``slq_operator`` has the C layout (no GPU needed)."""

import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from matfree_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "slq.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(slq_\w+)\(", text, re.M)))


def test_header_declares_the_bound_functions():
    assert declared_functions() == sorted(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libslq.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r"\bT (slq_\w+)", out))
    assert set(declared_functions()) <= exported


def test_status_codes_match_header():
    text = open(HEADER).read()
    for name in ["SLQ_OK", "SLQ_EINVAL", "SLQ_EINVAL_K", "SLQ_ENOMEM", "SLQ_EHIP",
                 "SLQ_EUNSUPPORTED", "SLQ_EADJOINT_K0"]:
        m = re.search(rf"{name}\s*=\s*(-?\d+)", text)
        assert m and int(m.group(1)) == getattr(_lib, name)


def test_operator_struct_layout_matches_c():
    src = r"""
#include <stddef.h>
#include <stdio.h>
#include "slq.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(slq_operator),
         offsetof(slq_operator, kind), offsetof(slq_operator, precision),
         offsetof(slq_operator, n), offsetof(slq_operator, matrix),
         offsetof(slq_operator, points), offsetof(slq_operator, dim),
         offsetof(slq_operator, storage), offsetof(slq_operator, lengthscale),
         offsetof(slq_operator, noise), offsetof(slq_operator, packed),
         offsetof(slq_operator, tile_scales));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), c, "-o", exe],
                       check=True)
        got = [int(x) for x in subprocess.run([exe], capture_output=True, text=True).stdout.split()]
    S = _lib.OperatorStruct
    want = [ctypes.sizeof(S)] + [getattr(S, f).offset for f in
                                 ["kind", "precision", "n", "matrix", "points", "dim", "storage",
                                  "lengthscale", "noise", "packed", "tile_scales"]]
    assert got == want


def test_null_context_is_rejected_without_a_gpu():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libslq.so not built")
    lib = _lib.load()
    assert lib.slq_hessenberg(None, None, None, 1, 1, None, None, None, None) == _lib.SLQ_EINVAL
    assert lib.slq_logdet(None, None, None, 1, 1, None, None, None, None) == _lib.SLQ_EINVAL
    assert lib.slq_workspace_bytes(None, 0, None, 1, 1, None) == _lib.SLQ_EINVAL
    assert lib.slq_workspace_growth(None, 0, None, 1, 1, None) == _lib.SLQ_EINVAL
    assert lib.slq_last_error(None) == b""
