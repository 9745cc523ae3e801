// Multisite tensor plan (src/siteTensor_init.c:609-800).
#include "sitetensor.hpp"

#include <algorithm>
#include <stdexcept>

namespace t3 {

namespace {

void indexize3(int* id, QN qn, const Symsecs* ss) {
    const QN n0 = ss[0].nrSecs, n1 = ss[1].nrSecs;
    id[0] = (int)(qn % n0);
    id[1] = (int)((qn / n0) % n1);
    id[2] = (int)(qn / (n0 * n1));
}

QN qntypize3(const int* id, const Symsecs* ss) {
    return (QN)id[0] + (QN)id[1] * ss[0].nrSecs + (QN)id[2] * ss[0].nrSecs * ss[1].nrSecs;
}

int sector_of(const Irrep& irrep, const Symsecs& ss) {
    for (int i = 0; i < ss.nrSecs; ++i)
        if (ss.irreps[i] == irrep) return i;
    return -1;
}

}  // namespace

MakesitePlan plan_makesite(const MakesiteInput& in) {
    const int ns = in.nrsites;
    if (ns < 2 || ns > STEPSPECS_MSITES || in.innersite < 0 || in.innersite >= ns)
        throw std::invalid_argument("makesite: needs 2-4 sites and an inner site");
    MakesitePlan P;
    P.nrsites = ns;
    P.innersite = in.innersite;
    const int I = in.innersite;
    // connections: every outer site shares one bond with the inner site
    // (init_md, src/siteTensor_init.c:192-260)
    int outer_of_leg[3] = {-1, -1, -1};
    for (int s = 0; s < ns; ++s) {
        if (s == I) continue;
        int found = 0;
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; k += 2)
                if (in.new_ss[I][j].bond == in.new_ss[s][k].bond) {
                    P.steps.push_back(MakesiteStep{s, j, k});
                    outer_of_leg[j] = s;
                    ++found;
                }
        if (found != 1) throw std::invalid_argument("makesite: outer site not bound to the inner site");
    }
    std::sort(P.steps.begin(), P.steps.end(),
              [](const MakesiteStep& a, const MakesiteStep& b) { return a.inner_leg < b.inner_leg; });
    // site order of the legs after the in-place contractions
    int key[3];
    for (int j = 0; j < 3; ++j) key[j] = outer_of_leg[j] >= 0 ? outer_of_leg[j] : I;
    std::stable_sort(P.perm, P.perm + 3, [&](int a, int b) { return key[a] < key[b]; });
    P.permuted = !(P.perm[0] == 0 && P.perm[1] == 1 && P.perm[2] == 2);

    P.site_arena_off.assign(ns + 1, 0);
    for (int s = 0; s < ns; ++s) P.site_arena_off[s + 1] = P.site_arena_off[s] + in.sites[s]->size();
    P.site_arena_size = P.site_arena_off[ns];
    const int nb = (int)in.beginblock.size() - 1;
    P.out_size = in.beginblock[nb];

    for (int sb = 0; sb < nb; ++sb) {
        MakesiteBlock B;
        B.out_off = in.beginblock[sb];
        B.out_size = in.beginblock[sb + 1] - in.beginblock[sb];
        bool valid = B.out_size > 0;
        // init_makeinfo (src/siteTensor_init.c:609-649)
        for (int s = 0; s < ns && valid; ++s) {
            int nids[3], oids[3];
            indexize3(nids, in.qnumbers[(size_t)sb * ns + s], in.new_ss[s]);
            for (int k = 0; k < 3; ++k) {
                oids[k] = sector_of(in.new_ss[s][k].irreps[nids[k]], in.old_ss[s][k]);
                if (oids[k] < 0) valid = false;
            }
            if (!valid) break;
            const QN oqn = qntypize3(oids, in.old_ss[s]);
            const SiteTensor& T = *in.sites[s];
            const auto it = std::lower_bound(T.qnumbers.begin(), T.qnumbers.end(), oqn);
            if (it == T.qnumbers.end() || *it != oqn) {
                valid = false;
                break;
            }
            const int csb = (int)(it - T.qnumbers.begin());
            if (T.block_size(csb) == 0) {
                valid = false;
                break;
            }
            B.site_off[s] = P.site_arena_off[s] + T.beginblock[csb];
            for (int k = 0; k < 3; ++k) B.odim[s][k] = in.old_ss[s][k].dims[oids[k]];
        }
        if (!valid) continue;
        // the product of the external leg dims is the multisite block size
        int64_t ext = 1;
        for (int j = 0; j < 3; ++j) {
            if (outer_of_leg[j] < 0) {
                ext *= B.odim[I][j];
            } else {
                const int s = outer_of_leg[j];
                int k = 0;
                for (const auto& st : P.steps)
                    if (st.site == s) k = st.site_leg;
                if (B.odim[s][k] != B.odim[I][j])
                    throw std::runtime_error("makesite: internal bond dims differ");
                ext *= (int64_t)B.odim[s][1] * B.odim[s][k == 0 ? 2 : 0];
            }
        }
        if (ext != B.out_size) throw std::runtime_error("makesite: block size mismatch");
        P.blocks.push_back(B);
    }
    return P;
}

}  // namespace t3
