"""The host-side argument checks of the C ABI (matfree_amd/csrc/slq_check.h:
the operator descriptor, the size limits, the packed-storage sizes) built on
their own with AddressSanitizer and UndefinedBehaviorSanitizer (any report is
fatal) and driven with adversarial inputs: NULL and partial descriptors,
out-of-range enums and dimensions, NaN / zero / negative lengthscales, and
sizes whose products would overflow 64-bit arithmetic.  No GPU needed."""

import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <string>
#include "slq_check.h"

using namespace slq;

static int fails = 0;
static void expect(const char* what, int got, int want) {
  if (got != want) {
    std::printf("FAIL %s: got %d want %d\n", what, got, want);
    ++fails;
  }
}

int main() {
  std::string msg;
  const float dummy = 0.f;
  const int64_t big = std::numeric_limits<int64_t>::max();
  expect("null descriptor", operator_status(nullptr, &msg), SLQ_EINVAL);

  slq_operator rbf{};
  rbf.kind = SLQ_OP_RBF;
  rbf.precision = SLQ_PREC_BF16X2;
  rbf.n = 1000;
  rbf.points = &dummy;
  rbf.dim = 3;
  rbf.lengthscale = 0.5;
  expect("valid rbf", operator_status(&rbf, &msg), SLQ_OK);
  for (int64_t n : {int64_t(0), int64_t(-5), kMaxDim + 1, big, std::numeric_limits<int64_t>::min()}) {
    slq_operator o = rbf;
    o.n = n;
    expect("rbf n", operator_status(&o, &msg), SLQ_EINVAL);
  }
  for (double l : {0.0, -1.0, std::nan(""), -std::numeric_limits<double>::infinity()}) {
    slq_operator o = rbf;
    o.lengthscale = l;
    expect("rbf lengthscale", operator_status(&o, &msg), SLQ_EINVAL);
  }
  for (int d : {0, -1, 5, 1 << 30}) {
    slq_operator o = rbf;
    o.dim = d;
    expect("rbf dim", operator_status(&o, &msg), SLQ_EINVAL);
  }
  {
    slq_operator o = rbf;
    o.dim = 4;
    expect("bf16 rbf dim 4", operator_status(&o, &msg), SLQ_EUNSUPPORTED);
    o.precision = SLQ_PREC_FP32;
    expect("fp32 rbf dim 4", operator_status(&o, &msg), SLQ_OK);
    o.points = nullptr;
    expect("rbf without points", operator_status(&o, &msg), SLQ_EINVAL);
  }
  for (int p : {-1, 4, 1 << 30, std::numeric_limits<int>::min()}) {
    slq_operator o = rbf;
    o.precision = p;
    expect("precision", operator_status(&o, &msg), SLQ_EUNSUPPORTED);
  }
  for (int kind : {-1, 2, std::numeric_limits<int>::max()}) {
    slq_operator o = rbf;
    o.kind = kind;
    expect("kind", operator_status(&o, &msg), SLQ_EINVAL);
  }

  slq_operator dense{};
  dense.kind = SLQ_OP_DENSE;
  dense.precision = SLQ_PREC_FP32;
  dense.n = kMaxDim;
  dense.matrix = &dummy;
  expect("dense at the limit", operator_status(&dense, &msg), SLQ_OK);
  dense.matrix = nullptr;
  expect("dense without matrix", operator_status(&dense, &msg), SLQ_EINVAL);
  for (int st : {-1, 5, 1 << 20}) {
    slq_operator o = dense;
    o.storage = st;
    o.matrix = &dummy;
    expect("storage", operator_status(&o, &msg), SLQ_EINVAL);
  }
  {
    slq_operator o = dense;
    o.storage = SLQ_STORE_FP8X2;
    o.packed = &dummy;
    expect("fp8 without scales", operator_status(&o, &msg), SLQ_EINVAL);
    o.tile_scales = &dummy;
    expect("fp8 with scales", operator_status(&o, &msg), SLQ_OK);
    o.storage = SLQ_STORE_BF16X2;
    o.tile_scales = nullptr;
    expect("bf16x2 packed", operator_status(&o, &msg), SLQ_OK);
    o.packed = nullptr;
    expect("packed without planes", operator_status(&o, &msg), SLQ_EINVAL);
  }

  expect("footprint small", footprint_status(128, 32768, 64, &msg), SLQ_OK);
  expect("footprint k < 0", footprint_status(1, 10, -3, &msg), SLQ_OK);
  expect("footprint P max", footprint_status(kMaxCount + 1, 10, 1, &msg), SLQ_EINVAL);
  expect("footprint P big", footprint_status(big, 10, 1, &msg), SLQ_EINVAL);
  expect("footprint product", footprint_status(kMaxCount, kMaxDim, kMaxDim, &msg), SLQ_EINVAL);
  expect("footprint k big", footprint_status(1, kMaxDim, big, &msg), SLQ_EINVAL);

  int64_t bytes = -1, scales = -1;
  expect("packed n 0", packed_layout(0, SLQ_STORE_BF16, &bytes, &scales), false);
  expect("packed n big", packed_layout(big, SLQ_STORE_FP8, &bytes, &scales), false);
  expect("packed fp32", packed_layout(100, SLQ_STORE_FP32, &bytes, &scales), false);
  expect("packed storage 9", packed_layout(100, 9, &bytes, &scales), false);
  expect("packed at the limit", packed_layout(kMaxDim, SLQ_STORE_FP8X2, &bytes, &scales), true);
  const int64_t nt = kMaxDim / kPackTile;
  expect("packed bytes", bytes == 2 * nt * nt * kTileFp8Bytes, true);
  expect("packed scales", scales == 2 * nt * nt, true);
  expect("packed n 70", packed_layout(70, SLQ_STORE_BF16X2, &bytes, &scales), true);
  expect("packed n 70 bytes", bytes == 2 * 9 * 2048 && scales == 0, true);

  std::printf("%s %d\n", fails ? "FAILED" : "OK", fails);
  return fails ? 1 : 0;
}
"""


def test_host_checks_under_asan_and_ubsan():
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "checks.cpp")
        exe = os.path.join(d, "checks")
        open(src, "w").write(DRIVER)
        subprocess.run([gxx, "-std=c++17", "-O1", "-g", "-Wall", "-Werror",
                        "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                        "-fno-omit-frame-pointer",
                        "-I", os.path.join(ROOT, "matfree_amd", "csrc"), src, "-o", exe],
                       check=True)
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
                   UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
        run = subprocess.run([exe], capture_output=True, text=True, env=env)
    print(run.stdout, run.stderr)
    assert run.returncode == 0, run.stdout + run.stderr
    assert run.stdout.strip().endswith("OK 0")
    assert "runtime error" not in run.stderr and "AddressSanitizer" not in run.stderr
