// Device copies of renormalised-operator arenas made on the device, kept
// after their host copy is handed out, and found again from an operator
// descriptor that points into that host copy: the next update or Heff
// set-up on those operators copies them on the device instead of sending
// them over the host link again.  No counterpart in the reference (its
// operators live in host memory only).
#pragma once
#include <cstdint>
#include <memory>

#include "hip_util.hpp"
#include "../host/tensors.hpp"

namespace t3 {

struct ResidentRef {
    std::shared_ptr<DevBuf<double>> buf;  // held for the duration of a use
    int64_t offset = 0;                   // of the operators' first element
    const double* ptr() const { return buf ? buf->get() + offset : nullptr; }
    explicit operator bool() const { return (bool)buf; }
};

// Records `dev` as the device copy of the host arena [host, host + n) on
// the current device.  Registered copies beyond the byte cap
// (T3NS_RESIDENT_GB, default 32; 0 turns residency off) are dropped oldest
// first.
void resident_put(const double* host, int64_t n, std::shared_ptr<DevBuf<double>> dev);
// Forgets the copy of the arena starting at `host` (its host copy is about
// to go).
void resident_drop(const double* host);
// The device copy holding operators `r` (built from a descriptor without a
// host copy: r.tel_src per operator, r.op_offset) in the same layout on the
// current device, or an empty reference.
ResidentRef resident_find(const ROperators& r);

}  // namespace t3
