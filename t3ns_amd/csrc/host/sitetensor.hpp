// Site-tensor operations around the Heff solve: building the multisite
// tensor from one-site tensors (src/siteTensor_init.c:609-800).  The plan is
// made on the host, the contractions run on the device
// (dev/sitetensor_dev.hip).
#pragma once
#include <cstdint>
#include <vector>

#include "model.hpp"
#include "sym.hpp"
#include "tensors.hpp"

namespace t3 {

// Everything makesiteTensor reads.
struct MakesiteInput {
    int nrsites = 0;
    int innersite = -1;
    Symsecs new_ss[STEPSPECS_MSITES][3];  // multisite sectors (internal bonds at unit dims)
    Symsecs old_ss[STEPSPECS_MSITES][3];  // sectors of the one-site tensors
    std::vector<QN> qnumbers;             // multisite: nrblocks * nrsites
    std::vector<int64_t> beginblock;      // multisite: nrblocks + 1
    const SiteTensor* sites[STEPSPECS_MSITES] = {};
};

// The outer site contracted into leg `inner_leg` of the inner site through
// its own leg `site_leg` (0 or 2).
struct MakesiteStep {
    int site, inner_leg, site_leg;
};

// One multisite block whose one-site blocks all exist: offsets of the
// inner and outer blocks in the concatenated one-site arena, the one-site
// leg dims, and the output offset.
struct MakesiteBlock {
    int64_t out_off, out_size;
    int64_t site_off[STEPSPECS_MSITES];
    int odim[STEPSPECS_MSITES][3];
};

struct MakesitePlan {
    int nrsites = 0, innersite = -1;
    std::vector<MakesiteStep> steps;  // in the order of the inner legs
    // After the in-place contractions the inner legs hold (inner external
    // leg or outer site's external legs); `perm` orders them by site as the
    // multisite layout requires (identity when already in order).
    int perm[3] = {0, 1, 2};
    bool permuted = false;
    std::vector<int64_t> site_arena_off;  // per site, start in the arena
    int64_t site_arena_size = 0;
    int64_t out_size = 0;
    std::vector<MakesiteBlock> blocks;
};

MakesitePlan plan_makesite(const MakesiteInput& in);

}  // namespace t3
