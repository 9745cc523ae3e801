// Autotuner of the effective-Hamiltonian work lists (see heff_tune.cpp).
#pragma once
#include <string>

#include "dev/heff_dev.hpp"

namespace t3 {

struct TuneReport {
    HeffTuning best;
    double ms = 0;           // matvec time under `best`
    double baseline_ms = 0;  // under the tuning the object had (0 on a cache hit)
    int trials = 0;          // tunings timed (0: taken from the cache)
    bool cached = false;
    std::string key;         // block-size histogram signature
};

// Block-size histogram signature of a device Heff (with its precision and
// object kind): the key of the tuning cache.
std::string heff_signature(const HeffDevice& dev);

// Times matvecs (x -> y, device arrays of dev.size(), on stream s) under at
// most `max_trials` tunings, leaves `dev` built for the fastest and records
// it for the signature; a signature seen before takes the recorded tuning
// without timing.  `in` is the input the device object was built from.
TuneReport autotune_heff(HeffDevice& dev, const HeffInput& in, hipStream_t s, const double* x,
                         double* y, int max_trials);

void clear_tune_cache();

}  // namespace t3
