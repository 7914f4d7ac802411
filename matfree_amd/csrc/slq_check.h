// Host-side argument checks of the C ABI that need no device: the operator
// descriptor, the size limits and the packed-storage sizes.  Plain C++ (no
// HIP), so that they can also be built on their own with the host sanitizers
// (tests/test_host_checks_sanitized.py).
#pragma once

#include <cstdint>
#include <string>

#include "../../include/slq.h"

namespace slq {

// Largest dimension n and probe / vector count P accepted.  Sizes derived from
// them (n^2 matrix elements, P n (k + 8) floats of scratch) then stay far
// below 2^63, so no size computation of the library can overflow.
constexpr int64_t kMaxDim = int64_t(1) << 30;
constexpr int64_t kMaxCount = int64_t(1) << 31;
// Largest device footprint (bytes) a request may imply; beyond any GPU, so
// only requests whose sizes could not be represented are refused here (a
// request that merely does not fit fails its allocation with SLQ_ENOMEM).
constexpr double kMaxFootprintBytes = 1e15;

// Packed storage (slq_storage): planes and bytes of one 32 x 32 tile.
constexpr int kPackTile = 32;
constexpr int64_t kTileBf16Bytes = 2048;
constexpr int64_t kTileFp8Bytes = 1024;
constexpr int storage_planes(int storage) {
  return (storage == SLQ_STORE_BF16X2 || storage == SLQ_STORE_FP8X2) ? 2 : 1;
}
constexpr bool storage_fp8(int storage) {
  return storage == SLQ_STORE_FP8X2 || storage == SLQ_STORE_FP8;
}

// SLQ_OK, or a status with *msg set: the operator descriptor.
inline int operator_status(const slq_operator* op, std::string* msg) {
  if (!op) {
    *msg = "operator descriptor is NULL";
    return SLQ_EINVAL;
  }
  if (op->n <= 0) {
    *msg = "operator dimension n must be positive";
    return SLQ_EINVAL;
  }
  if (op->n > kMaxDim) {
    *msg = "operator dimension n exceeds " + std::to_string(kMaxDim);
    return SLQ_EINVAL;
  }
  if (op->precision != SLQ_PREC_FP32 && op->precision != SLQ_PREC_BF16 &&
      op->precision != SLQ_PREC_BF16X2 && op->precision != SLQ_PREC_BF16X3) {
    *msg = "precision " + std::to_string(op->precision) + " is not available in this build";
    return SLQ_EUNSUPPORTED;
  }
  if (op->kind == SLQ_OP_DENSE) {
    if (op->storage == SLQ_STORE_FP32) {
      if (!op->matrix) {
        *msg = "dense operator without a matrix";
        return SLQ_EINVAL;
      }
    } else if (op->storage >= SLQ_STORE_BF16X2 && op->storage <= SLQ_STORE_FP8) {
      if (!op->packed) {
        *msg = "packed dense operator without its planes";
        return SLQ_EINVAL;
      }
      if (storage_fp8(op->storage) && !op->tile_scales) {
        *msg = "FP8 dense operator without its tile scales";
        return SLQ_EINVAL;
      }
    } else {
      *msg = "unknown matrix storage " + std::to_string(op->storage);
      return SLQ_EINVAL;
    }
  } else if (op->kind == SLQ_OP_RBF) {
    if (!op->points) {
      *msg = "RBF operator without points";
      return SLQ_EINVAL;
    }
    if (op->dim < 1 || op->dim > 4) {
      *msg = "RBF point dimension must be in 1..4";
      return SLQ_EINVAL;
    }
    // also rejects NaN; an infinite lengthscale gives a constant kernel
    if (!(op->lengthscale > 0.0)) {
      *msg = "RBF lengthscale must be > 0";
      return SLQ_EINVAL;
    }
    if (op->precision != SLQ_PREC_FP32 && op->dim > 3) {
      *msg = "BF16 RBF kernels support point dimension <= 3";
      return SLQ_EUNSUPPORTED;
    }
  } else {
    *msg = "unknown operator kind " + std::to_string(op->kind);
    return SLQ_EINVAL;
  }
  return SLQ_OK;
}

// SLQ_OK, or SLQ_EINVAL with *msg set: P vectors of dimension n over depth
// k (k = 0 for a single pass), P > 0 checked by the caller.
inline int footprint_status(int64_t P, int64_t n, int64_t k, std::string* msg) {
  if (P > kMaxCount) {
    *msg = "number of vectors P exceeds " + std::to_string(kMaxCount);
    return SLQ_EINVAL;
  }
  // in doubles: a bound, and no product can overflow
  const double bytes = 4.0 * double(P) * double(n) * (double(k < 0 ? 0 : k) + 8.0);
  if (!(bytes <= kMaxFootprintBytes)) {
    *msg = "P = " + std::to_string(P) + ", n = " + std::to_string(n) + ", k = " +
           std::to_string(k) + " implies more than 1e15 bytes of device memory";
    return SLQ_EINVAL;
  }
  return SLQ_OK;
}

// Sizes of the packed form of an n x n matrix (slq_packed_size); false for
// an invalid n or storage.
inline bool packed_layout(int64_t n, int storage, int64_t* bytes, int64_t* scale_count) {
  if (n <= 0 || n > kMaxDim || storage <= SLQ_STORE_FP32 || storage > SLQ_STORE_FP8) return false;
  const int64_t nt = (n + kPackTile - 1) / kPackTile;
  const int planes = storage_planes(storage);
  const bool fp8 = storage_fp8(storage);
  *bytes = planes * nt * nt * (fp8 ? kTileFp8Bytes : kTileBf16Bytes);
  *scale_count = fp8 ? planes * nt * nt : 0;
  return true;
}

}  // namespace slq
