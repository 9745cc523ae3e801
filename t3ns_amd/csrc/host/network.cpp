// Tree tensor network topology and sweep scheduler; see network.hpp.
#include "network.hpp"

#include <cassert>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace t3 {

namespace {
bool is_end_marker(const std::string& line) {
    std::string s;
    for (char c : line)
        if (!isspace((unsigned char)c)) s += c;
    return s == "&END" || s == "/END" || s == "/";
}
std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}
}  // namespace

Network Network::from_text(const std::string& text) {
    // Header "KEY = value" lines up to an end marker, then the site->orbital
    // line ('*' marks a branching site), optionally the sweep line, then one
    // "from to" pair per bond (src/network.c:207-354).
    std::istringstream in(text);
    std::string line;
    Network n;
    int sweeplength = -1;
    while (std::getline(in, line)) {
        if (is_end_marker(line)) break;
        auto eq = line.find('=');
        if (eq == std::string::npos) continue;
        std::string key = trim(line.substr(0, eq));
        int val = std::stoi(trim(line.substr(eq + 1)));
        if (key == "NR_SITES") n.sites = val;
        else if (key == "NR_PHYS_SITES") n.psites = val;
        else if (key == "NR_BONDS") n.nr_bonds = val;
        else if (key == "SWEEP_LENGTH") sweeplength = val;
    }
    if (!std::getline(in, line)) throw std::runtime_error("netw: missing site line");
    {
        std::istringstream ls(line);
        std::string tok;
        while (ls >> tok) n.sitetoorb.push_back(tok == "*" ? -1 : std::stoi(tok));
    }
    if ((int)n.sitetoorb.size() != n.sites)
        throw std::runtime_error("netw: wrong number of sites in the site line");
    int np = 0;
    for (int o : n.sitetoorb) np += o >= 0;
    if (np != n.psites) throw std::runtime_error("netw: wrong number of physical sites");
    while (std::getline(in, line))
        if (is_end_marker(line)) break;
    if (sweeplength != -1) {
        std::getline(in, line);
        std::istringstream ls(line);
        int v;
        while (ls >> v) n.sweep.push_back(v);
        if ((int)n.sweep.size() != sweeplength) throw std::runtime_error("netw: bad sweep line");
        while (std::getline(in, line))
            if (is_end_marker(line)) break;
    }
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        int a, b;
        if (!(ls >> a >> b)) continue;
        if (a < -1 || a >= n.sites || b < -1 || b >= n.sites)
            throw std::runtime_error("netw: illegal site in bond list");
        n.bonds.push_back({a, b});
    }
    if ((int)n.bonds.size() != n.nr_bonds) throw std::runtime_error("netw: bond count mismatch");
    n.finalize();
    return n;
}

Network Network::from_file(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open network file " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return from_text(ss.str());
}

Network Network::from_bonds(int nsites, const std::vector<int>& s2o,
                            const std::vector<std::array<int, 2>>& bonds) {
    return from_bonds(nsites, s2o, bonds, {});
}

Network Network::from_bonds(int nsites, const std::vector<int>& s2o,
                            const std::vector<std::array<int, 2>>& bonds,
                            const std::vector<int>& sweep) {
    Network n;
    if ((int)s2o.size() != nsites) throw std::runtime_error("network: sitetoorb of the wrong length");
    for (const auto& b : bonds)
        for (int k = 0; k < 2; ++k)
            if (b[k] < -1 || b[k] >= nsites) throw std::runtime_error("network: illegal site in bond list");
    for (int v : sweep)
        if (v < 0 || v >= (int)bonds.size()) throw std::runtime_error("network: illegal bond in sweep");
    n.sweep = sweep;
    n.sites = nsites;
    n.sitetoorb = s2o;
    n.psites = 0;
    for (int o : s2o) n.psites += o >= 0;
    n.bonds = bonds;
    n.nr_bonds = (int)bonds.size();
    n.finalize();
    return n;
}

void Network::finalize() {
    // Validity: one outgoing vacuum bond, leg counts (src/network.c:13-52).
    int ends = 0;
    for (auto& b : bonds) ends += b[1] == -1;
    if (ends != 1) throw std::runtime_error("network must have exactly one ending bond");
    std::vector<int> legs(sites, 0);
    for (auto& b : bonds)
        if (b[0] != -1 && b[1] != -1) {
            ++legs[b[0]];
            ++legs[b[1]];
        }
    for (int i = 0; i < sites; ++i) {
        bool okp = (legs[i] == 1 || legs[i] == 2) && is_psite(i);
        bool okb = legs[i] <= 3 && !is_psite(i);
        if (!okp && !okb) throw std::runtime_error("network site with illegal leg count");
    }
    // Number of physical sites left of every bond (src/network.c:77-102).
    nr_left_psites.assign(nr_bonds, 0);
    std::vector<int> temp(nr_bonds, 0);
    for (int bond = 0; bond < nr_bonds; ++bond) {
        if (bonds[bond][0] != -1) continue;
        std::fill(temp.begin(), temp.end(), 0);
        int curr = bond, ns;
        while ((ns = bonds[curr][1]) != -1) {
            const int prev = curr;
            for (curr = 0; curr < nr_bonds; ++curr)
                if (bonds[curr][0] == ns) break;
            temp[curr] += temp[prev] + (nr_left_psites[curr] == 0) * (int)is_psite(ns);
            nr_left_psites[curr] += temp[curr];
        }
    }
    // Orbital order left/right of every bond (src/network.c:104-183).
    order_psites.assign(nr_bonds, std::vector<int>(psites, 0));
    for (int bond = 0; bond < nr_bonds; ++bond) {
        const int site = bonds[bond][0];
        if (site == -1) continue;
        int b[3];
        bonds_of_site(site, b);
        if (is_psite(site)) {
            int i;
            for (i = 0; i < nr_left_psites[b[0]]; ++i) order_psites[bond][i] = order_psites[b[0]][i];
            order_psites[bond][i] = sitetoorb[site];
        } else {
            for (int i = 0; i < nr_left_psites[b[0]]; ++i)
                order_psites[bond][i] = order_psites[b[0]][i];
            for (int i = 0; i < nr_left_psites[b[1]]; ++i)
                order_psites[bond][nr_left_psites[b[0]] + i] = order_psites[b[1]][i];
        }
    }
    for (int bond = nr_bonds - 1; bond >= 0; --bond) {
        const int site = bonds[bond][1];
        if (site == -1) continue;
        int b[3];
        bonds_of_site(site, b);
        if (is_psite(site)) {
            for (int i = nr_left_psites[b[2]]; i < psites; ++i)
                order_psites[bond][i - 1] = order_psites[b[2]][i];
            order_psites[bond][psites - 1] = sitetoorb[site];
        } else {
            const int b1 = b[0] == bond ? b[1] : b[0];
            const int b2 = b[2];
            for (int i = 0; i < nr_left_psites[b1]; ++i)
                order_psites[bond][nr_left_psites[bond] + i] = order_psites[b1][i];
            for (int i = nr_left_psites[b2]; i < psites; ++i)
                order_psites[bond][nr_left_psites[bond] + nr_left_psites[b1] + i -
                                   nr_left_psites[b2]] = order_psites[b2][i];
        }
    }
    if (sweep.empty()) make_simplesweep(false);
    curr_state_ = 0;
}

const int* Network::order_psites_of(int bond, bool is_left) const {
    return &order_psites[bond][is_left ? 0 : nr_left_psites[bond]];
}

void Network::bonds_of_site(int site, int* out) const {
    out[0] = out[1] = out[2] = -1;
    int i;
    for (i = 0; i < nr_bonds; ++i)
        if (bonds[i][1] == site) {
            out[0] = i;
            break;
        }
    if (is_psite(site)) {
        out[1] = 2 * nr_bonds + sitetoorb[site];
    } else {
        for (++i; i < nr_bonds; ++i)
            if (bonds[i][1] == site) {
                out[1] = i;
                break;
            }
    }
    for (i = 0; i < nr_bonds; ++i)
        if (bonds[i][0] == site) {
            out[2] = i;
            break;
        }
}

int Network::common_bond(int s1, int s2) const {
    int b1[3], b2[3];
    bonds_of_site(s1, b1);
    bonds_of_site(s2, b2);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (b1[i] == b2[j]) return b1[i];
    return -1;
}

bool Network::is_dmrg_bond(int bond) const {
    return bonds[bond][0] >= 0 && bonds[bond][1] >= 0 && is_psite(bonds[bond][0]) &&
           is_psite(bonds[bond][1]);
}

int Network::outgoing_bond() const {
    for (int i = 0; i < nr_bonds; ++i)
        if (bonds[i][1] == -1) return i;
    return -1;
}

// ------------------------------------------------------------------ sweeps
void Network::recursive_stepping(bool inclborder, int* id) {
    // Depth-first walk against, then with, the flow (src/network.c:748-800).
    int site = sweep[*id - 1];
    int b[3];
    bool flag = true;
    do {
        bonds_of_site(site, b);
        site = bonds[b[0]][0];
        if (site == -1) {
            *id -= !inclborder;
            site = sweep[*id - 1];
            flag = false;
        } else {
            sweep[(*id)++] = site;
        }
    } while (flag && is_psite(site));
    if (flag) {
        bonds_of_site(site, b);
        sweep[(*id)++] = bonds[b[1]][0];
        recursive_stepping(inclborder, id);
        sweep[(*id)++] = bonds[b[0]][0];
        recursive_stepping(inclborder, id);
    }
    while (flag || is_psite(site)) {
        flag = false;
        bonds_of_site(site, b);
        site = bonds[b[2]][1];
        if (site == -1) {
            --*id;
            break;
        } else {
            sweep[(*id)++] = site;
        }
    }
}

void Network::make_simplesweep(bool inclborder) {
    // src/network.c:802-831
    sweep.assign(nr_bonds * 4, 0);
    int len = 0;
    for (int i = 0; i < nr_bonds; ++i)
        if (bonds[i][1] == -1) {
            sweep[len++] = bonds[i][0];
            break;
        }
    if (len != 1 || !is_psite(sweep[0]))
        throw std::runtime_error("outgoing bond must end on a physical site");
    recursive_stepping(inclborder, &len);
    sweep.resize(len);
}

void Network::move_forward_state(const StepSpecs* specs, int* state) const {
    const int swl = (int)sweep.size();
    bool flag = false;
    while (!flag) {
        flag = true;
        const int next_site = sweep[(*state + 1) % swl];
        for (int i = 0; i < specs->nr_sites_opt; ++i)
            if (specs->sites_opt[i] == next_site) {
                flag = false;
                break;
            }
        if (!flag) ++*state;
    }
}

bool Network::get_sites_to_opt(int maxsites, StepSpecs* specs, int* state) const {
    // Returns true when the sweep is finished (src/network.c:549-605).
    const int swl = (int)sweep.size();
    int* so = specs->sites_opt;
    if (*state >= swl) {
        *state = 0;
        return true;
    }
    specs->nr_sites_opt = 0;
    so[specs->nr_sites_opt++] = sweep[*state];
    if (maxsites == 1) {
        ++*state;
        return false;
    }
    so[specs->nr_sites_opt++] = sweep[(*state + 1) % swl];
    const int cbond = common_bond(so[0], so[1]);
    if (!(maxsites == 2 || is_dmrg_bond(cbond))) {
        const int branch = is_psite(so[0]) ? so[1] : so[0];
        if (maxsites == 4) {
            int b[3];
            bonds_of_site(branch, b);
            so[0] = bonds[b[0]][0];
            so[1] = bonds[b[1]][0];
            so[2] = bonds[b[2]][1];
            so[3] = branch;
            specs->nr_sites_opt = 4;
        }
        if (maxsites == 3) {
            const int nextsite = sweep[(*state + 2) % swl];
            bool add = true;
            for (int i = 0; i < specs->nr_sites_opt; ++i)
                if (nextsite == so[i] || !is_psite(nextsite)) {
                    add = false;
                    break;
                }
            if (add) so[specs->nr_sites_opt++] = nextsite;
        }
    }
    move_forward_state(specs, state);
    return false;
}

void Network::get_bonds_involved(StepSpecs* specs) const {
    // External virtual bonds of the multisite object (src/network.c:614-649).
    int b[STEPSPECS_MSITES][3];
    for (int i = 0; i < specs->nr_sites_opt; ++i) bonds_of_site(specs->sites_opt[i], b[i]);
    specs->nr_bonds_opt = 0;
    const int n3 = 3 * specs->nr_sites_opt;
    for (int i = 0; i < n3; ++i) {
        const int ib = b[i / 3][i % 3];
        if (ib >= 2 * nr_bonds) continue;
        int j;
        for (j = 0; j < n3; ++j)
            if (ib == b[j / 3][j % 3] && i != j) break;
        if (j != n3) continue;
        assert(specs->nr_bonds_opt < STEPSPECS_MBONDS);
        specs->bonds_opt[specs->nr_bonds_opt++] = ib;
    }
    int* bo = specs->bonds_opt;
    if (bo[0] > bo[1]) std::swap(bo[0], bo[1]);
    if (specs->nr_bonds_opt == STEPSPECS_MBONDS && bo[1] > bo[2]) std::swap(bo[1], bo[2]);
    if (bo[0] > bo[1]) std::swap(bo[0], bo[1]);
}

void Network::get_common_with_next(int maxsites, StepSpecs* specs, int next_state) const {
    for (int i = 0; i < specs->nr_sites_opt; ++i) specs->common_next[i] = 0;
    const int swl = (int)sweep.size();
    if (maxsites == 1) {
        const int nsite = sweep[next_state % swl];
        const int cbond = common_bond(specs->sites_opt[0], nsite);
        for (int i = 0; i < specs->nr_bonds_opt; ++i)
            if (specs->bonds_opt[i] == cbond) {
                specs->common_next[0] = i;
                return;
            }
        throw std::runtime_error("no common bond with next step");
    }
    StepSpecs nxt;
    if (get_sites_to_opt(maxsites, &nxt, &next_state)) {
        const int last = sweep[0];
        for (int i = 0; i < specs->nr_sites_opt; ++i)
            if (last == specs->sites_opt[i]) {
                specs->common_next[i] = 1;
                break;
            }
    } else {
        for (int i = 0; i < specs->nr_sites_opt; ++i)
            for (int j = 0; j < nxt.nr_sites_opt; ++j)
                if (specs->sites_opt[i] == nxt.sites_opt[j]) {
                    specs->common_next[i] = 1;
                    break;
                }
    }
}

void Network::set_nCenter(StepSpecs* specs) const {
    if (specs->nr_sites_opt == 1) {
        const int site = specs->sites_opt[0];
        const auto& mb = bonds[specs->bonds_opt[specs->common_next[0]]];
        specs->nCenter = mb[0] == site ? mb[1] : mb[0];
    } else {
        for (int i = 0; i < specs->nr_sites_opt; ++i)
            if (specs->common_next[i]) {
                specs->nCenter = specs->sites_opt[i];
                if (is_psite(specs->nCenter)) break;
            }
    }
}

bool Network::next_opt_step(int maxsites, StepSpecs* specs) {
    if (maxsites > STEPSPECS_MSITES) maxsites = STEPSPECS_MSITES;
    if (get_sites_to_opt(maxsites, specs, &curr_state_)) return false;
    get_bonds_involved(specs);
    get_common_with_next(maxsites, specs, curr_state_);
    set_nCenter(specs);
    return true;
}

Network make_chain_network(int L) {
    std::vector<int> s2o(L);
    std::vector<std::array<int, 2>> b;
    for (int i = 0; i < L; ++i) s2o[i] = i;
    b.push_back({-1, 0});
    for (int i = 1; i < L; ++i) b.push_back({i - 1, i});
    b.push_back({L - 1, -1});
    return Network::from_bonds(L, s2o, b);
}

}  // namespace t3
