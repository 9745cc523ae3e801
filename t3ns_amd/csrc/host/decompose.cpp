// SVD split plan and singular-value selection
// (src/siteTensor_decompose.c:587-1444).
#include "decompose.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace t3 {

namespace {

void indexize3(int* id, QN qn, const Symsecs* ss) {
    const QN n0 = ss[0].nrSecs, n1 = ss[1].nrSecs;
    id[0] = (int)(qn % n0);
    id[1] = (int)((qn / n0) % n1);
    id[2] = (int)(qn / (n0 * n1));
}

int64_t full_size(const Symsecs* ss, QN qn) {
    int id[3];
    indexize3(id, qn, ss);
    return (int64_t)ss[0].dims[id[0]] * ss[1].dims[id[1]] * ss[2].dims[id[2]];
}

}  // namespace

SplitPlan plan_split(const SplitInput& in) {
    const int ns = in.nrsites;
    if (ns < 2 || in.siteV < 0 || in.siteV >= ns) throw std::invalid_argument("split: bad site");
    SplitPlan P;
    P.id_siteV = in.siteV;
    // the cut bond: V must share exactly one bond with the other sites
    // (good_site_to_split 1074-1118, init_svddata 862-887)
    int links = 0;
    for (int i = 0; i < ns; ++i) {
        if (i == in.siteV) continue;
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k)
                if (in.ss[in.siteV][k].bond == in.ss[i][j].bond) {
                    ++links;
                    P.id_csite = i;
                    P.id_cbond = j;
                    P.id_bond = k;
                }
    }
    if (links != 1) throw std::invalid_argument("split: the site is not cut off by one bond");
    const int nb = (int)in.beginblock.size() - 1;
    // U and V structures: unique sorted quantum-number tuples (init_splitted_tens)
    P.U_nrsites = ns - 1;
    std::vector<std::vector<QN>> ut;
    std::vector<QN> vt;
    for (int b = 0; b < nb; ++b) {
        std::vector<QN> t;
        for (int i = 0; i < ns; ++i)
            if (i != in.siteV) t.push_back(in.qnumbers[(size_t)b * ns + i]);
        ut.push_back(t);
        vt.push_back(in.qnumbers[(size_t)b * ns + in.siteV]);
    }
    // tuples compare from the last site to the first (sort_qn[n], src/sort.c)
    auto less_tuple = [](const std::vector<QN>& a, const std::vector<QN>& b) {
        for (int i = (int)a.size() - 1; i >= 0; --i)
            if (a[i] != b[i]) return a[i] < b[i];
        return false;
    };
    std::vector<std::vector<QN>> us = ut;
    std::sort(us.begin(), us.end(), less_tuple);
    us.erase(std::unique(us.begin(), us.end()), us.end());
    std::vector<QN> vs = vt;
    std::sort(vs.begin(), vs.end());
    vs.erase(std::unique(vs.begin(), vs.end()), vs.end());
    for (const auto& t : us) P.U_qn.insert(P.U_qn.end(), t.begin(), t.end());
    P.V_qn = vs;

    // sectors of the cut bond (make_r_count_svdinfos)
    const Symsecs* sV = in.ss[in.siteV];
    const Symsecs* sC = in.ss[P.id_csite];
    P.nrSss = sC[P.id_cbond].nrSecs;
    if (sV[P.id_bond].nrSecs != P.nrSss) throw std::invalid_argument("split: cut bond sectors differ");
    P.sec.resize(P.nrSss);
    QN divV = 1;
    for (int i = 0; i < P.id_bond; ++i) divV *= sV[i].nrSecs;
    const QN modV = sV[P.id_bond].nrSecs;
    for (int b = 0; b < nb; ++b) {
        const int s = (int)((vt[b] / divV) % modV);
        P.sec[s].idA.push_back(b);
    }
    const int csiteU = P.id_csite - (P.id_csite > in.siteV);
    QN divC = 1;
    for (int i = 0; i < P.id_cbond; ++i) divC *= sC[i].nrSecs;
    const QN modC = sC[P.id_cbond].nrSecs;
    // symsecs of U's sites in U order
    std::vector<const Symsecs*> ussU;
    for (int i = 0; i < ns; ++i)
        if (i != in.siteV) ussU.push_back(in.ss[i]);
    for (size_t u = 0; u < us.size(); ++u) {
        const int s = (int)((us[u][csiteU] / divC) % modC);
        SplitSector& S = P.sec[s];
        if (S.Mstart.empty()) S.Mstart.push_back(0);
        int64_t sz = 1;
        for (int i = 0; i < P.U_nrsites; ++i) sz *= full_size(ussU[i], us[u][i]);
        S.idU.push_back((int)u);
        S.Mstart.push_back(S.Mstart.back() + sz);
    }
    for (size_t v = 0; v < vs.size(); ++v) {
        const int s = (int)((vs[v] / divV) % modV);
        SplitSector& S = P.sec[s];
        if (S.Nstart.empty()) S.Nstart.push_back(0);
        S.idV.push_back((int)v);
        S.Nstart.push_back(S.Nstart.back() + full_size(sV, vs[v]));
    }
    for (auto& S : P.sec) {
        if (S.Mstart.empty()) S.Mstart.push_back(0);
        if (S.Nstart.empty()) S.Nstart.push_back(0);
        S.M = S.Mstart.back();
        S.N = S.Nstart.back();
        S.dimS = std::min(S.M, S.N);
    }
    // A blocks: sector matrix positions (SVD_copy_to_mem)
    P.ablocks.resize(nb);
    for (int b = 0; b < nb; ++b) {
        SplitPlan::ABlock& A = P.ablocks[b];
        const int s = (int)((vt[b] / divV) % modV);
        const SplitSector& S = P.sec[s];
        A.sector = s;
        int idU = -1, idV = -1;
        for (size_t k = 0; k < S.idU.size(); ++k)
            if (us[S.idU[k]] == ut[b]) {
                idU = (int)k;
                break;
            }
        for (size_t k = 0; k < S.idV.size(); ++k)
            if (vs[S.idV[k]] == vt[b]) {
                idV = (int)k;
                break;
            }
        if (idU < 0 || idV < 0) throw std::runtime_error("split: block not found");
        A.row0 = S.Mstart[idU];
        A.col0 = S.Nstart[idV];
        A.d[0] = A.d[1] = A.d[2] = 1;
        int cnt = 0;
        for (int i = 0; i < ns; ++i) {
            const int64_t sz = full_size(in.ss[i], in.qnumbers[(size_t)b * ns + i]);
            if (i == in.siteV) {
                A.d[1] = sz;
                cnt = 2;
            } else {
                A.d[cnt] *= sz;
            }
        }
        if (A.d[0] * A.d[1] * A.d[2] != in.beginblock[b + 1] - in.beginblock[b])
            throw std::runtime_error("split: block size mismatch");
    }
    return P;
}

std::vector<int> select_singular_values(const std::vector<std::vector<double>>& sing,
                                        const std::vector<int>& mult, const SvalSelect& sel,
                                        SelectRes* res) {
    if (sel.truncType != 'E' && sel.truncType != 'N')
        throw std::invalid_argument("selectS: truncation type must be E or N");
    auto vn = [](double s) {
        const double w = s * s;
        return w < 1e-20 ? 0.0 : -w * std::log(w);
    };
    const int nsec = (int)sing.size();
    res->norm[0] = res->entropy[0] = 0;
    std::vector<double> all;
    std::vector<int> mul, orig;
    for (int i = 0; i < nsec; ++i) {
        const double mf = 1 / std::sqrt((double)mult[i]);
        double n = 0, e = 0;
        for (double s : sing[i]) {
            n += (s * mf) * (s * mf);
            e += vn(s * mf);
            all.push_back(s);
            mul.push_back(mult[i]);
            orig.push_back(i);
        }
        res->norm[0] += mult[i] * n;
        res->entropy[0] += mult[i] * e;
    }
    // ascending, ties in original order (a stable sort), taken from the top
    const int total = (int)all.size();
    std::vector<int> idx(total);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return all[a] < all[b]; });
    const int runupto = std::min(sel.maxD, total);
    const int minimalto = std::min(sel.minD, total);
    std::vector<int> kept(nsec, 0);
    res->norm[1] = res->entropy[1] = 0;
    int ss = 0;
    auto take = [&](int k) {
        const double s = all[k] / std::sqrt((double)mul[k]);
        res->norm[1] += mul[k] * s * s;
        res->entropy[1] += mul[k] * vn(s);
        ++kept[orig[k]];
    };
    for (; ss < minimalto; ++ss) take(idx[total - ss - 1]);
    for (; ss < runupto; ++ss) {
        take(idx[total - ss - 1]);
        const bool ok = sel.truncType == 'N'
                            ? res->norm[0] - res->norm[1] < sel.truncerr
                            : res->entropy[0] - res->entropy[1] / res->norm[1] - std::log(res->norm[1]) <
                                  sel.truncerr;
        if (ok) break;
    }
    res->entropy[1] = res->entropy[1] / res->norm[1] + std::log(res->norm[1]);
    return kept;
}

}  // namespace t3

namespace t3 {

QRPlan plan_qr(const QRInput& in) {
    QRPlan P;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (in.ssA[i].bond == in.ssC[j].bond) {
                P.bondA = i;
                P.bondC = j;
            }
    if (P.bondA < 0) throw std::invalid_argument("qr: the tensors share no bond");
    const Symsecs* ss = in.ssA;
    const int b = P.bondA;
    const int nsec = ss[b].nrSecs;
    P.sec.resize(nsec);
    const SiteTensor& A = *in.A;
    P.ablocks.resize(A.nrblocks);
    // blocks grouped by their sector of the bond, in block order (a stable
    // sort of the indices, sort_indices 38-46 with makestart 77-97)
    for (int k = 0; k < A.nrblocks; ++k) {
        int id[3];
        indexize3(id, A.qnumbers[k], ss);
        const int s = id[b];
        QRSector& S = P.sec[s];
        int64_t before = 1, after = 1;
        for (int i = 0; i < b; ++i) before *= ss[i].dims[id[i]];
        for (int i = b + 1; i < 3; ++i) after *= ss[i].dims[id[i]];
        P.ablocks[k] = QRPlan::ABlock{s, before, after};
        if (A.block_size(k) == 0) continue;
        S.blocks.push_back(k);
        S.rowstart.push_back(S.M);
        S.M += before * after;
    }
    for (int s = 0; s < nsec; ++s) {
        QRSector& S = P.sec[s];
        S.N = S.blocks.empty() ? 0 : ss[b].dims[s];
        if (S.N == 0) S.M = 0;
        S.K = std::min(S.M, S.N);
    }
    return P;
}

}  // namespace t3
