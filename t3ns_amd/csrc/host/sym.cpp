// Symmetry algebra; see sym.hpp for the reference files this follows.
#include <cstdint>
#include <unordered_map>
#include "sym.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <vector>
#include <cctype>

namespace t3 {

namespace {
const char* kSymNames[] = {"Z2", "U1", "SU2", "C1", "Ci", "C2", "Cs",
                           "D2", "C2v", "C2h", "D2h", "SENIORITY"};
// Irrep counts of the real abelian point groups C1..D2h
// (src/symmetry_pg.c:26).
const int kPgCount[8] = {1, 2, 2, 2, 4, 4, 4, 8};
// psi4 irrep names per point group (src/symmetry_pg.c:28-37).
const char* kPgNames[8][8] = {
    {"A"},
    {"Ag", "Au"},
    {"A", "B"},
    {"A'", "A''"},
    {"A", "B1", "B2", "B3"},
    {"A1", "A2", "B1", "B2"},
    {"Ag", "Bg", "Au", "Bu"},
    {"Ag", "B1g", "B2g", "B3g", "Au", "B1u", "B2u", "B3u"}};
// Molpro -> psi4 irrep order (src/symmetry_pg.c:50-56, 85-89).
const int kMolproToPsi4[5][8] = {{0}, {0, 1}, {0, 3, 2, 1}, {0, 2, 3, 1},
                                 {0, 7, 6, 1, 5, 2, 3, 4}};
const int kPgToTable[8] = {0, 1, 1, 1, 2, 3, 3, 4};

bool ieq(const std::string& a, const char* b) {
    size_t n = std::strlen(b);
    if (a.size() != n) return false;
    for (size_t i = 0; i < n; ++i)
        if (std::tolower((unsigned char)a[i]) != std::tolower((unsigned char)b[i])) return false;
    return true;
}

// Factorials in extended precision.  n! for n <= 400 fits in the x87 long
// double range; the cumulative product keeps ~19 significant digits.
struct FactTable {
    long double f[401];
    FactTable() {
        f[0] = 1.0L;
        for (int i = 1; i <= 400; ++i) f[i] = f[i - 1] * (long double)i;
    }
};
const FactTable& facts() {
    static FactTable t;
    return t;
}

inline bool triangle_bad(int a, int b, int c) {
    return ((a + b + c) & 1) || (a + b < c) || (std::abs(a - b) > c);
}

// Triangle coefficient Delta(abc)^2 with doubled arguments.
inline long double delta2(int a, int b, int c) {
    const long double* f = facts().f;
    return f[(a + b - c) / 2] * f[(a - b + c) / 2] * f[(-a + b + c) / 2] / f[(a + b + c) / 2 + 1];
}

inline double bracket(int twoj) { return std::sqrt((double)(twoj + 1)); }
inline int sign_if(bool neg) { return neg ? -1 : 1; }
}  // namespace

const char* symstring(int sg) {
    if (sg < 0 || sg > SENIORITY) return "INVALID";
    return kSymNames[sg];
}

int symgroup_from_string(const std::string& s) {
    for (int i = 0; i <= SENIORITY; ++i)
        if (ieq(s, kSymNames[i])) return i;
    return -1;
}

int pg_nr_irreps(int sg) { return kPgCount[sg - C1]; }

int pg_irrep_from_string(int sg, const std::string& s) {
    const int pg = sg - C1;
    for (int i = 0; i < kPgCount[pg]; ++i)
        if (ieq(s, kPgNames[pg][i])) return i;
    return -1;
}

int fcidump_to_psi4(int fcidump_irrep0, int sg) {
    return kMolproToPsi4[kPgToTable[sg - C1]][fcidump_irrep0];
}

void tensprod_irrep(int* min_irrep, int* nr_irreps, int* step, int irrep1, int irrep2,
                    int sign, int sg) {
    switch (sg) {
        case Z2:
            *nr_irreps = 1;
            *step = 1;
            *min_irrep = (irrep1 + irrep2) % 2;
            return;
        case U1:
            *min_irrep = irrep1 + sign * irrep2;
            *nr_irreps = 1;
            *step = 1;
            return;
        case SU2: {
            const int hi = irrep1 + irrep2;
            *min_irrep = std::abs(irrep1 - irrep2);
            *nr_irreps = (hi - *min_irrep) / 2 + 1;
            *step = 2;
            return;
        }
        case SENIORITY:
            throw std::runtime_error("SENIORITY symmetry is not supported");
        default:
            // Real abelian point groups: irreps combine by XOR.
            *nr_irreps = 1;
            *step = 1;
            *min_irrep = irrep1 ^ irrep2;
            return;
    }
}

int max_irrep_bound(int max1, int max2, int sg) {
    switch (sg) {
        case Z2: return 2;
        case U1:
        case SU2: return max1 + max2 + 1;
        default: return kPgCount[sg - C1];
    }
}

double wigner6j(int a, int b, int c, int d, int e, int f) {
    // {a b c; d e f} (doubled spins).  Racah: the four triads are
    // (a b c), (a e f), (d b f), (d e c).
    if (triangle_bad(a, b, c) || triangle_bad(a, e, f) || triangle_bad(d, b, f) ||
        triangle_bad(d, e, c))
        return 0.0;
    const int t1 = (a + b + c) / 2, t2 = (a + e + f) / 2, t3 = (d + b + f) / 2,
              t4 = (d + e + c) / 2;
    const int u1 = (a + b + d + e) / 2, u2 = (b + c + e + f) / 2, u3 = (a + c + d + f) / 2;
    int tmin = t1;
    if (t2 > tmin) tmin = t2;
    if (t3 > tmin) tmin = t3;
    if (t4 > tmin) tmin = t4;
    int tmax = u1;
    if (u2 < tmax) tmax = u2;
    if (u3 < tmax) tmax = u3;
    if (tmax < tmin) return 0.0;
    const long double* F = facts().f;
    long double sum = 0.0L;
    for (int t = tmin; t <= tmax; ++t) {
        long double term = F[t + 1] / (F[t - t1] * F[t - t2] * F[t - t3] * F[t - t4] *
                                       F[u1 - t] * F[u2 - t] * F[u3 - t]);
        sum += (t & 1) ? -term : term;
    }
    const long double norm =
        sqrtl(delta2(a, b, c) * delta2(a, e, f) * delta2(d, b, f) * delta2(d, e, c));
    return (double)(sum * norm);
}

static double wigner9j_eval(int j1, int j2, int j3, int j4, int j5, int j6, int j7, int j8,
                            int j9);

double wigner9j(int j1, int j2, int j3, int j4, int j5, int j6, int j7, int j8, int j9) {
    // Memoised: a Heff plan evaluates the same few symbols for millions of
    // (block, block, MPO) triples.  Doubled spins below 128 pack into one key.
    const int v[9] = {j1, j2, j3, j4, j5, j6, j7, j8, j9};
    uint64_t key = 0;
    bool packable = true;
    for (int x : v) {
        packable = packable && x >= 0 && x < 128;
        key = (key << 7) | (uint64_t)(x & 127);
    }
    if (!packable) return wigner9j_eval(j1, j2, j3, j4, j5, j6, j7, j8, j9);
    thread_local std::unordered_map<uint64_t, double> memo;
    const auto it = memo.find(key);
    if (it != memo.end()) return it->second;
    const double w = wigner9j_eval(j1, j2, j3, j4, j5, j6, j7, j8, j9);
    memo.emplace(key, w);
    return w;
}

static double wigner9j_eval(int j1, int j2, int j3, int j4, int j5, int j6, int j7, int j8,
                            int j9) {
    // {j1 j2 j3; j4 j5 j6; j7 j8 j9} = sum_x (-1)^(2x) (2x+1)
    //      {j1 j4 j7; j8 j9 x} {j2 j5 j8; j4 x j6} {j3 j6 j9; x j1 j2}
    if (triangle_bad(j1, j2, j3) || triangle_bad(j4, j5, j6) || triangle_bad(j7, j8, j9) ||
        triangle_bad(j1, j4, j7) || triangle_bad(j2, j5, j8) || triangle_bad(j3, j6, j9))
        return 0.0;
    int lo = std::abs(j1 - j9);
    if (std::abs(j8 - j4) > lo) lo = std::abs(j8 - j4);
    if (std::abs(j2 - j6) > lo) lo = std::abs(j2 - j6);
    int hi = j1 + j9;
    if (j8 + j4 < hi) hi = j8 + j4;
    if (j2 + j6 < hi) hi = j2 + j6;
    long double sum = 0.0L;
    for (int x = lo; x <= hi; x += 2) {
        const long double w = (long double)(x + 1) * wigner6j(j1, j4, j7, j8, j9, x) *
                              wigner6j(j2, j5, j8, j4, x, j6) *
                              wigner6j(j3, j6, j9, x, j1, j2);
        sum += w;
    }
    return (double)((lo & 1) ? -sum : sum);
}

// ---------------------------------------------------------------- prefactors
// Each prefactor is a product over the symmetry groups; only Z2 (fermionic
// signs) and SU2 (recoupling) contribute.

double prefactor_pAppend(const int* const (*irr)[3], int is_left, const SymInfo& s) {
    // src/symmetry_z2.c:54-60, src/symmetry_su2.c:75-93
    double p = 1;
    for (int y = 0; y < s.nrSyms; ++y) {
        int v[3][3];
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) v[j][k] = irr[j][k][y];
        if (s.sgs[y] == Z2) {
            if (is_left)
                p *= sign_if(v[1][1] && v[2][0]);
            else
                p *= sign_if((v[2][1] * v[1][2] + v[2][1]) % 2);
        } else if (s.sgs[y] == SU2) {
            if (is_left) {
                p *= bracket(v[0][2]) * bracket(v[1][2]) * bracket(v[2][2]) *
                     wigner9j(v[0][0], v[1][0], v[2][0], v[0][1], v[1][1], v[2][1], v[0][2],
                              v[1][2], v[2][2]);
            } else {
                p *= sign_if((v[0][1] + v[1][1] + v[2][1]) % 4) * bracket(v[0][0]) *
                     bracket(v[1][0]) * bracket(v[2][2]) *
                     wigner9j(v[0][0], v[1][0], v[2][0], v[0][2], v[1][2], v[2][2], v[0][1],
                              v[1][1], v[2][1]);
            }
        }
    }
    return p;
}

double prefactor_adjoint(const int* const* irr, char c, const SymInfo& s) {
    // src/symmetry_z2.c:62-84
    double p = 1;
    for (int y = 0; y < s.nrSyms; ++y) {
        if (s.sgs[y] != Z2) continue;
        switch (c) {
            case 'c': p *= sign_if(irr[2][y]); break;
            case '1': p *= sign_if(irr[1][y]); break;
            case '2': p *= sign_if(irr[0][y]); break;
            case '3': break;
            default: throw std::runtime_error("prefactor_adjoint: bad option");
        }
    }
    return p;
}

double prefactor_pUpdate(const int* const (*irr)[3], int is_left, const SymInfo& s) {
    // src/symmetries.c:213-235, src/symmetry_z2.c:86-103
    double p = 1;
    for (int y = 0; y < s.nrSyms; ++y) {
        if (s.sgs[y] != Z2 || is_left) continue;
        // ket(i) is irr[0][1], bra(i) is irr[1][1] in the 7-vector layout
        p *= sign_if((irr[0][1][y] + irr[1][1][y]) % 2);
    }
    return p;
}

double prefactor_mirror_coupling(const int* const* irr, const SymInfo& s) {
    // src/symmetry_z2.c:155-159, src/symmetry_su2.c:70-73
    double p = 1;
    for (int y = 0; y < s.nrSyms; ++y) {
        if (s.sgs[y] == Z2)
            p *= sign_if((irr[0][y] + irr[1][y] * irr[2][y]) % 2);
        else if (s.sgs[y] == SU2)
            p *= bracket(irr[0][y]);
    }
    return p;
}

double prefactor_bUpdate(const int* const (*irr)[3], int u, const SymInfo& s) {
    // src/symmetry_z2.c:105-153, src/symmetry_su2.c:110-120
    double p = 1;
    for (int y = 0; y < s.nrSyms; ++y) {
        int v[3][3];
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) v[j][k] = irr[j][k][y];
        if (s.sgs[y] == Z2) {
            switch (u) {
                case 0: p *= sign_if((v[2][1] * v[1][2] + v[1][0]) % 2); break;
                case 1: p *= sign_if((v[0][1] * v[2][2] + v[0][0]) % 2); break;
                case 2: p *= sign_if(v[1][1] && v[0][2]); break;
                default: throw std::runtime_error("prefactor_bUpdate: bad case");
            }
        } else if (s.sgs[y] == SU2) {
            const int sg = sign_if((v[2][0] + v[2][1] + v[2][2] + v[u][0] + v[u][1] + v[u][2]) % 4);
            p *= sg * bracket(v[u][0]) * bracket(v[u][1]) * bracket(v[2][2]) *
                 wigner9j(v[0][0], v[0][1], v[0][2], v[1][0], v[1][1], v[1][2], v[2][0], v[2][1],
                          v[2][2]);
        }
    }
    return p;
}

double prefactor_add_P_operator(const int* const (*irr)[3], int isleft, const SymInfo& s) {
    // src/symmetry_z2.c:161-209
    double p = 1;
    for (int y = 0; y < s.nrSyms; ++y) {
        if (s.sgs[y] != Z2 || isleft) continue;
        p *= sign_if((irr[1][2][y] + irr[1][0][y]) % 2);
    }
    return p;
}

double prefactor_combine_MPOs(const int* const (*irr)[3], const int* const* mpo,
                              const SymInfo& s, int isdmrg, int e) {
    // src/symmetry_z2.c:211-228, src/symmetry_su2.c:95-108
    double p = 1;
    for (int y = 0; y < s.nrSyms; ++y) {
        if (s.sgs[y] == Z2) {
            if (isdmrg)
                p *= sign_if((irr[1][e][y] + irr[0][e][y] * mpo[1][y]) % 2);
            else
                p *= sign_if((irr[1][1][y] * mpo[0][y] + irr[1][2][y] + irr[0][2][y] * mpo[2][y]) % 2);
        } else if (s.sgs[y] == SU2) {
            if (isdmrg) {
                p *= (1.0 / bracket(irr[0][e][y])) * (1.0 / bracket(irr[1][e][y])) *
                     sign_if((irr[0][e][y] + irr[1][e][y] + mpo[1][y]) % 4);
            } else {
                p *= sign_if((irr[0][2][y] + irr[1][2][y] + mpo[2][y]) % 4) * bracket(mpo[2][y]) *
                     wigner9j(irr[0][0][y], irr[1][0][y], mpo[0][y], irr[0][1][y], irr[1][1][y],
                              mpo[1][y], irr[0][2][y], irr[1][2][y], mpo[2][y]);
            }
        }
    }
    return p;
}

bool need_multiplicity(const SymInfo& s) {
    for (int i = 0; i < s.nrSyms; ++i)
        if (s.sgs[i] == SU2) return true;
    return false;
}

int multiplicity(const SymInfo& s, const int* irrep) {
    int r = 1;
    for (int i = 0; i < s.nrSyms; ++i)
        if (s.sgs[i] == SU2) r *= irrep[i] + 1;
    return r;
}

}  // namespace t3
