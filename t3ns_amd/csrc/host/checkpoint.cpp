// Checkpoints in the reference's HDF5 layout; see checkpoint.hpp.  Group,
// attribute and dataset names, element types and shapes follow the
// reference's writers (src/io_to_disk.c:35-437, src/hamiltonian.c:327-345,
// src/hamiltonian_qc.c:640-654, src/hamiltonian_nn_hubbard.c:261-270) so
// that the files are interchangeable; the code is this library's own.
#include "checkpoint.hpp"

#include <dlfcn.h>

#include <cstdlib>
#include <mutex>
#include <stdexcept>

#if __has_include(<hdf5.h>)
#include <hdf5.h>
#define T3NS_HDF5_HEADERS 1
#endif

namespace t3 {

#ifdef T3NS_HDF5_HEADERS
namespace {

// The HDF5 entry points this module uses, resolved from the shared
// library at first use.
struct H5Api {
    decltype(&H5open) open = nullptr;
    decltype(&H5Eset_auto2) eset_auto = nullptr;
    decltype(&H5Fcreate) fcreate = nullptr;
    decltype(&H5Fopen) fopen = nullptr;
    decltype(&H5Fclose) fclose = nullptr;
    decltype(&H5Gcreate2) gcreate = nullptr;
    decltype(&H5Gopen2) gopen = nullptr;
    decltype(&H5Gclose) gclose = nullptr;
    decltype(&H5Lexists) lexists = nullptr;
    decltype(&H5Aexists) aexists = nullptr;
    decltype(&H5Screate_simple) screate = nullptr;
    decltype(&H5Sget_simple_extent_npoints) snpoints = nullptr;
    decltype(&H5Sclose) sclose = nullptr;
    decltype(&H5Acreate2) acreate = nullptr;
    decltype(&H5Awrite) awrite = nullptr;
    decltype(&H5Aopen) aopen = nullptr;
    decltype(&H5Aread) aread = nullptr;
    decltype(&H5Aget_space) aspace = nullptr;
    decltype(&H5Aclose) aclose = nullptr;
    decltype(&H5Dcreate2) dcreate = nullptr;
    decltype(&H5Dwrite) dwrite = nullptr;
    decltype(&H5Dopen2) dopen = nullptr;
    decltype(&H5Dread) dread = nullptr;
    decltype(&H5Dget_space) dspace = nullptr;
    decltype(&H5Dclose) dclose = nullptr;
    // file types of the reference (int, double, EL_TYPE double, QN_TYPE
    // int64; include/io_to_disk.h) and the matching memory types
    hid_t i32le = -1, f64le = -1, i64le = -1;
    hid_t n_int = -1, n_double = -1, n_llong = -1;
};

// File access flags without the header's macros (they call H5open and
// H5check_version, which are resolved at run time here): the values of
// H5F_ACC_RDONLY and H5F_ACC_TRUNC in H5Fpublic.h.
constexpr unsigned kAccRdonly = 0x0000u, kAccTrunc = 0x0002u;

std::once_flag g_h5_once;
H5Api g_h5;
std::string g_h5_error;

template <typename F>
void resolve(void* lib, const char* name, F& f) {
    f = reinterpret_cast<F>(dlsym(lib, name));
    if (!f) throw std::runtime_error(std::string("HDF5 library lacks ") + name);
}
hid_t type_global(void* lib, const char* name) {
    auto* p = reinterpret_cast<hid_t*>(dlsym(lib, name));
    if (!p) throw std::runtime_error(std::string("HDF5 library lacks ") + name);
    return *p;
}

const H5Api& h5() {
    std::call_once(g_h5_once, [] {
        try {
            const char* env = std::getenv("T3NS_HDF5_LIB");
            const char* names[] = {env, "libhdf5.so.103", "/opt/conda/lib/libhdf5.so.103",
                                   "libhdf5.so"};
            void* lib = nullptr;
            for (const char* n : names)
                if (n && (lib = dlopen(n, RTLD_NOW | RTLD_LOCAL))) break;
            if (!lib) throw std::runtime_error("HDF5 library not found (set T3NS_HDF5_LIB)");
            H5Api a;
            resolve(lib, "H5open", a.open);
            resolve(lib, "H5Eset_auto2", a.eset_auto);
            resolve(lib, "H5Fcreate", a.fcreate);
            resolve(lib, "H5Fopen", a.fopen);
            resolve(lib, "H5Fclose", a.fclose);
            resolve(lib, "H5Gcreate2", a.gcreate);
            resolve(lib, "H5Gopen2", a.gopen);
            resolve(lib, "H5Gclose", a.gclose);
            resolve(lib, "H5Lexists", a.lexists);
            resolve(lib, "H5Aexists", a.aexists);
            resolve(lib, "H5Screate_simple", a.screate);
            resolve(lib, "H5Sget_simple_extent_npoints", a.snpoints);
            resolve(lib, "H5Sclose", a.sclose);
            resolve(lib, "H5Acreate2", a.acreate);
            resolve(lib, "H5Awrite", a.awrite);
            resolve(lib, "H5Aopen", a.aopen);
            resolve(lib, "H5Aread", a.aread);
            resolve(lib, "H5Aget_space", a.aspace);
            resolve(lib, "H5Aclose", a.aclose);
            resolve(lib, "H5Dcreate2", a.dcreate);
            resolve(lib, "H5Dwrite", a.dwrite);
            resolve(lib, "H5Dopen2", a.dopen);
            resolve(lib, "H5Dread", a.dread);
            resolve(lib, "H5Dget_space", a.dspace);
            resolve(lib, "H5Dclose", a.dclose);
            if (a.open() < 0) throw std::runtime_error("H5open failed");
            a.eset_auto(H5E_DEFAULT, nullptr, nullptr);  // errors come back as codes
            a.i32le = type_global(lib, "H5T_STD_I32LE_g");
            a.f64le = type_global(lib, "H5T_IEEE_F64LE_g");
            a.i64le = type_global(lib, "H5T_STD_I64LE_g");
            a.n_int = type_global(lib, "H5T_NATIVE_INT_g");
            a.n_double = type_global(lib, "H5T_NATIVE_DOUBLE_g");
            a.n_llong = type_global(lib, "H5T_NATIVE_LLONG_g");
            g_h5 = a;
        } catch (const std::exception& e) {
            g_h5_error = e.what();
        }
    });
    if (!g_h5.open) throw std::runtime_error(g_h5_error);
    return g_h5;
}

void ok(herr_t rc, const std::string& what) {
    if (rc < 0) throw std::runtime_error("HDF5: " + what + " failed");
}
hid_t okid(hid_t id, const std::string& what) {
    if (id < 0) throw std::runtime_error("HDF5: " + what + " failed");
    return id;
}

// Element kinds: file type and memory type.
enum Kind { K_INT, K_DOUBLE, K_QN };
hid_t ftype(Kind k) {
    const H5Api& a = h5();
    return k == K_INT ? a.i32le : (k == K_DOUBLE ? a.f64le : a.i64le);
}
hid_t mtype(Kind k) {
    const H5Api& a = h5();
    return k == K_INT ? a.n_int : (k == K_DOUBLE ? a.n_double : a.n_llong);
}

// One-dimensional attribute / dataset of `n` elements; empty ones are not
// written (as the reference skips them).
void put_attr(hid_t g, const char* name, const void* v, hsize_t n, Kind k) {
    if (!v || n == 0) return;
    const H5Api& a = h5();
    const hid_t sp = okid(a.screate(1, &n, nullptr), "dataspace");
    const hid_t at = okid(a.acreate(g, name, ftype(k), sp, H5P_DEFAULT, H5P_DEFAULT),
                          std::string("attribute ") + name);
    ok(a.awrite(at, mtype(k), v), std::string("attribute ") + name);
    a.aclose(at);
    a.sclose(sp);
}
void put_data(hid_t g, const char* name, const void* v, hsize_t n, Kind k) {
    if (!v || n == 0) return;
    const H5Api& a = h5();
    const hid_t sp = okid(a.screate(1, &n, nullptr), "dataspace");
    const hid_t ds = okid(a.dcreate(g, name, ftype(k), sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                          std::string("dataset ") + name);
    ok(a.dwrite(ds, mtype(k), H5S_ALL, H5S_ALL, H5P_DEFAULT, v), std::string("dataset ") + name);
    a.dclose(ds);
    a.sclose(sp);
}
template <typename T>
std::vector<T> get_attr(hid_t g, const char* name, Kind k) {
    const H5Api& a = h5();
    const hid_t at = okid(a.aopen(g, name, H5P_DEFAULT), std::string("attribute ") + name);
    const hid_t sp = a.aspace(at);
    const hssize_t n = a.snpoints(sp);
    std::vector<T> v(n > 0 ? (size_t)n : 0);
    if (n > 0) ok(a.aread(at, mtype(k), v.data()), std::string("attribute ") + name);
    a.sclose(sp);
    a.aclose(at);
    return v;
}
template <typename T>
T get_attr1(hid_t g, const char* name, Kind k) {
    std::vector<T> v = get_attr<T>(g, name, k);
    if (v.size() != 1) throw std::runtime_error(std::string("checkpoint: attribute ") + name +
                                                " is not a scalar");
    return v[0];
}
template <typename T>
std::vector<T> get_data(hid_t g, const char* name, Kind k, size_t expect) {
    const H5Api& a = h5();
    if (expect == 0 && a.lexists(g, name, H5P_DEFAULT) <= 0) return {};
    const hid_t ds = okid(a.dopen(g, name, H5P_DEFAULT), std::string("dataset ") + name);
    const hid_t sp = a.dspace(ds);
    const hssize_t n = a.snpoints(sp);
    a.sclose(sp);
    if (n < 0 || (size_t)n != expect) {
        a.dclose(ds);
        throw std::runtime_error(std::string("checkpoint: dataset ") + name + " has " +
                                 std::to_string(n) + " elements, expected " +
                                 std::to_string(expect));
    }
    std::vector<T> v((size_t)n);
    if (n > 0) ok(a.dread(ds, mtype(k), H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()),
                  std::string("dataset ") + name);
    a.dclose(ds);
    return v;
}

struct Group {  // closes on scope exit
    hid_t id;
    explicit Group(hid_t i) : id(i) {}
    ~Group() {
        if (id >= 0) h5().gclose(id);
    }
    Group(const Group&) = delete;
    Group& operator=(const Group&) = delete;
};
hid_t mkgroup(hid_t parent, const std::string& name) {
    return okid(h5().gcreate(parent, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                "group " + name);
}
hid_t opgroup(hid_t parent, const std::string& name) {
    return okid(h5().gopen(parent, name.c_str(), H5P_DEFAULT), "group " + name);
}

// struct symsecs (src/io_to_disk.c:35-51): irreps padded to MAX_SYMMETRIES.
void write_symsec(hid_t g, const Symsecs& s, int nmbr, char kind) {
    Group h(mkgroup(g, std::string("./") + kind + "_symsec_" + std::to_string(nmbr)));
    put_attr(h.id, "nrSecs", &s.nrSecs, 1, K_INT);
    put_attr(h.id, "bond", &s.bond, 1, K_INT);
    put_attr(h.id, "totaldims", &s.totaldims, 1, K_INT);
    put_data(h.id, "./dims", s.dims.data(), s.dims.size(), K_INT);
    std::vector<int> irr((size_t)s.nrSecs * MAX_SYMMETRIES, 0);
    for (int i = 0; i < s.nrSecs; ++i)
        for (int j = 0; j < MAX_SYMMETRIES; ++j) irr[(size_t)i * MAX_SYMMETRIES + j] = s.irreps[i][j];
    put_data(h.id, "./irreps", irr.data(), irr.size(), K_INT);
    put_data(h.id, "./fcidims", s.fcidims.data(), s.fcidims.size(), K_DOUBLE);
}
Symsecs read_symsec(hid_t g, int nmbr, char kind, int stride, int nrsyms) {
    Group h(opgroup(g, std::string("./") + kind + "_symsec_" + std::to_string(nmbr)));
    Symsecs s;
    s.nrSecs = get_attr1<int>(h.id, "nrSecs", K_INT);
    s.bond = get_attr1<int>(h.id, "bond", K_INT);
    s.totaldims = get_attr1<int>(h.id, "totaldims", K_INT);
    s.dims = get_data<int>(h.id, "./dims", K_INT, s.nrSecs);
    const std::vector<int> irr = get_data<int>(h.id, "./irreps", K_INT, (size_t)s.nrSecs * stride);
    s.irreps.assign(s.nrSecs, Irrep{});
    for (int i = 0; i < s.nrSecs; ++i)
        for (int j = 0; j < nrsyms; ++j) s.irreps[i][j] = irr[(size_t)i * stride + j];
    s.fcidims = get_data<double>(h.id, "./fcidims", K_DOUBLE, s.nrSecs);
    return s;
}

// struct sparseblocks (src/io_to_disk.c:143-164): int32 block offsets.
void write_blocks(hid_t g, const SiteBlocks& t, int nmbr) {
    Group h(mkgroup(g, "./block_" + std::to_string(nmbr)));
    const int nb = t.nrblocks();
    put_attr(h.id, "nrBlocks", &nb, 1, K_INT);
    if (nb == 0) return;
    std::vector<int> bb(t.beginblock.size());
    for (size_t i = 0; i < bb.size(); ++i) {
        if (t.beginblock[i] > INT32_MAX)
            throw std::runtime_error("checkpoint: a site tensor exceeds the format's int32 block offsets");
        bb[i] = (int)t.beginblock[i];
    }
    put_data(h.id, "./beginblock", bb.data(), bb.size(), K_INT);
    put_data(h.id, "./tel", t.tel.data(), (hsize_t)t.beginblock.back(), K_DOUBLE);
}

}  // namespace

void write_checkpoint(const std::string& path, const Context& c,
                      const std::vector<SiteBlocks>& tensors) {
    const H5Api& a = h5();
    const Network& n = c.netw;
    if ((int)tensors.size() != n.sites)
        throw std::invalid_argument("checkpoint: one site tensor per network site expected");
    const hid_t f = okid(a.fcreate(path.c_str(), kAccTrunc, H5P_DEFAULT, H5P_DEFAULT),
                         "create " + path);
    try {
        {  // /network (src/io_to_disk.c:376-393)
            Group g(mkgroup(f, "/network"));
            put_attr(g.id, "nr_bonds", &n.nr_bonds, 1, K_INT);
            std::vector<int> bonds;
            for (const auto& b : n.bonds) bonds.insert(bonds.end(), {b[0], b[1]});
            put_data(g.id, "./bonds", bonds.data(), bonds.size(), K_INT);
            put_attr(g.id, "psites", &n.psites, 1, K_INT);
            put_attr(g.id, "sites", &n.sites, 1, K_INT);
            put_data(g.id, "./sitetoorb", n.sitetoorb.data(), n.sitetoorb.size(), K_INT);
            const int sl = (int)n.sweep.size();
            put_attr(g.id, "sweeplength", &sl, 1, K_INT);
            put_data(g.id, "./sweep", n.sweep.data(), n.sweep.size(), K_INT);
        }
        {  // /bookkeeper (src/io_to_disk.c:87-108)
            Group g(mkgroup(f, "/bookkeeper"));
            const int mx = MAX_SYMMETRIES;
            put_attr(g.id, "nrSyms", &c.sym.nrSyms, 1, K_INT);
            put_attr(g.id, "Max_symmetries", &mx, 1, K_INT);
            put_attr(g.id, "sgs", c.sym.sgs, c.sym.nrSyms, K_INT);
            put_attr(g.id, "target_state", c.sym.target_state, c.sym.nrSyms, K_INT);
            const int nb = (int)c.v_symsecs.size();
            put_attr(g.id, "nr_bonds", &nb, 1, K_INT);
            for (int i = 0; i < nb; ++i) write_symsec(g.id, c.v_symsecs[i], i, 'v');
            const int np = (int)c.p_symsecs.size();
            put_attr(g.id, "psites", &np, 1, K_INT);
            for (int i = 0; i < np; ++i) write_symsec(g.id, c.p_symsecs[i], i, 'p');
        }
        {  // /hamiltonian (src/hamiltonian.c:327-345 and the models' writers)
            const HamData d = c.ham->ham_data();
            Group g(mkgroup(f, "/hamiltonian"));
            put_attr(g.id, "type", &d.type, 1, K_INT);
            Group h(mkgroup(g.id, "./hamiltonian_data"));
            if (d.type == HAM_QC) {
                put_attr(h.id, "norb", &d.norb, 1, K_INT);
                put_data(h.id, "./orbirrep", d.orbirrep.data(), d.orbirrep.size(), K_INT);
                put_data(h.id, "./Vijkl", d.Vijkl.data(), d.Vijkl.size(), K_DOUBLE);
                put_attr(h.id, "core_energy", &d.core_energy, 1, K_DOUBLE);
                put_attr(h.id, "su2", &d.su2, 1, K_INT);
                put_attr(h.id, "has_seniority", &d.has_seniority, 1, K_INT);
            } else {
                put_attr(h.id, "t", &d.t, 1, K_DOUBLE);
                put_attr(h.id, "U", &d.U, 1, K_DOUBLE);
                put_attr(h.id, "su2", &d.su2, 1, K_INT);
            }
        }
        {  // /T3NS (src/io_to_disk.c:194-225, 238-250)
            Group g(mkgroup(f, "/T3NS"));
            put_attr(g.id, "nrSites", &n.sites, 1, K_INT);
            for (int i = 0; i < n.sites; ++i) {
                const SiteBlocks& t = tensors[i];
                const int ns = (int)t.sites.size(), nb = t.nrblocks();
                if (nb < 0 || t.qnumbers.size() != (size_t)nb * ns ||
                    (int64_t)t.tel.size() < (nb >= 0 && !t.beginblock.empty() ? t.beginblock.back() : 0))
                    throw std::invalid_argument("checkpoint: inconsistent site tensor " +
                                                std::to_string(i));
                Group h(mkgroup(g.id, "./tensor_" + std::to_string(i)));
                put_attr(h.id, "nrsites", &ns, 1, K_INT);
                put_attr(h.id, "sites", t.sites.data(), ns, K_INT);
                put_attr(h.id, "nrblocks", &nb, 1, K_INT);
                put_data(h.id, "./qnumbers", t.qnumbers.data(), t.qnumbers.size(), K_QN);
                write_blocks(h.id, t, 0);
            }
        }
        {  // /rOps (src/io_to_disk.c:347-357): no operators stored
            Group g(mkgroup(f, "/rOps"));
            put_attr(g.id, "nrOps", &n.nr_bonds, 1, K_INT);
        }
    } catch (...) {
        a.fclose(f);
        throw;
    }
    ok(a.fclose(f), "close " + path);
}

CheckpointData read_checkpoint(const std::string& path) {
    const H5Api& a = h5();
    const hid_t f = okid(a.fopen(path.c_str(), kAccRdonly, H5P_DEFAULT), "open " + path);
    CheckpointData d;
    try {
        {  // /network (src/io_to_disk.c:395-419)
            Group g(opgroup(f, "/network"));
            const int nb = get_attr1<int>(g.id, "nr_bonds", K_INT);
            const std::vector<int> bonds = get_data<int>(g.id, "./bonds", K_INT, (size_t)2 * nb);
            const int sites = get_attr1<int>(g.id, "sites", K_INT);
            const std::vector<int> s2o = get_data<int>(g.id, "./sitetoorb", K_INT, sites);
            const int sl = get_attr1<int>(g.id, "sweeplength", K_INT);
            std::vector<int> sweep = get_data<int>(g.id, "./sweep", K_INT, sl);
            std::vector<std::array<int, 2>> bl(nb);
            for (int i = 0; i < nb; ++i) bl[i] = {bonds[2 * i], bonds[2 * i + 1]};
            d.netw = Network::from_bonds(sites, s2o, bl, sweep);
            if (get_attr1<int>(g.id, "psites", K_INT) != d.netw.psites)
                throw std::runtime_error("checkpoint: psites disagrees with sitetoorb");
        }
        {  // /bookkeeper (src/io_to_disk.c:110-141)
            Group g(opgroup(f, "/bookkeeper"));
            d.sym.nrSyms = get_attr1<int>(g.id, "nrSyms", K_INT);
            if (d.sym.nrSyms < 1 || d.sym.nrSyms > MAX_SYMMETRIES)
                throw std::runtime_error("checkpoint: " + std::to_string(d.sym.nrSyms) +
                                         " symmetries (this build holds at most " +
                                         std::to_string(MAX_SYMMETRIES) + ")");
            d.max_symmetries = get_attr1<int>(g.id, "Max_symmetries", K_INT);
            if (d.max_symmetries < d.sym.nrSyms) throw std::runtime_error("checkpoint: bad Max_symmetries");
            const std::vector<int> sgs = get_attr<int>(g.id, "sgs", K_INT);
            const std::vector<int> ts = get_attr<int>(g.id, "target_state", K_INT);
            if ((int)sgs.size() != d.sym.nrSyms || (int)ts.size() != d.sym.nrSyms)
                throw std::runtime_error("checkpoint: sgs / target_state of the wrong length");
            for (int i = 0; i < d.sym.nrSyms; ++i) {
                d.sym.sgs[i] = sgs[i];
                d.sym.target_state[i] = ts[i];
            }
            const int nb = get_attr1<int>(g.id, "nr_bonds", K_INT);
            for (int i = 0; i < nb; ++i)
                d.v_symsecs.push_back(read_symsec(g.id, i, 'v', d.max_symmetries, d.sym.nrSyms));
            const int np = get_attr1<int>(g.id, "psites", K_INT);
            for (int i = 0; i < np; ++i)
                d.p_symsecs.push_back(read_symsec(g.id, i, 'p', d.max_symmetries, d.sym.nrSyms));
        }
        {  // /hamiltonian
            Group g(opgroup(f, "/hamiltonian"));
            d.ham.type = get_attr1<int>(g.id, "type", K_INT);
            Group h(opgroup(g.id, "./hamiltonian_data"));
            if (d.ham.type == HAM_QC) {
                d.ham.norb = get_attr1<int>(h.id, "norb", K_INT);
                const size_t n = d.ham.norb;
                d.ham.orbirrep = get_data<int>(h.id, "./orbirrep", K_INT, n);
                d.ham.Vijkl = get_data<double>(h.id, "./Vijkl", K_DOUBLE, n * n * n * n);
                d.ham.core_energy = get_attr1<double>(h.id, "core_energy", K_DOUBLE);
                d.ham.su2 = get_attr1<int>(h.id, "su2", K_INT);
                d.ham.has_seniority = get_attr1<int>(h.id, "has_seniority", K_INT);
            } else if (d.ham.type == HAM_NN_HUBBARD) {
                d.ham.t = get_attr1<double>(h.id, "t", K_DOUBLE);
                d.ham.U = get_attr1<double>(h.id, "U", K_DOUBLE);
                d.ham.su2 = get_attr1<int>(h.id, "su2", K_INT);
            }
        }
        {  // /T3NS
            Group g(opgroup(f, "/T3NS"));
            const int ns = get_attr1<int>(g.id, "nrSites", K_INT);
            if (ns != d.netw.sites) throw std::runtime_error("checkpoint: nrSites disagrees with the network");
            for (int i = 0; i < ns; ++i) {
                Group h(opgroup(g.id, "./tensor_" + std::to_string(i)));
                SiteBlocks t;
                const int nrs = get_attr1<int>(h.id, "nrsites", K_INT);
                t.sites = get_attr<int>(h.id, "sites", K_INT);
                if ((int)t.sites.size() != nrs) throw std::runtime_error("checkpoint: bad sites attribute");
                const int nb = get_attr1<int>(h.id, "nrblocks", K_INT);
                t.qnumbers = get_data<int64_t>(h.id, "./qnumbers", K_QN, (size_t)nb * nrs);
                Group b(opgroup(h.id, "./block_0"));
                if (get_attr1<int>(b.id, "nrBlocks", K_INT) != nb)
                    throw std::runtime_error("checkpoint: nrBlocks disagrees with nrblocks");
                t.beginblock.assign(1, 0);
                if (nb > 0) {
                    const std::vector<int> bb = get_data<int>(b.id, "./beginblock", K_INT, (size_t)nb + 1);
                    t.beginblock.assign(bb.begin(), bb.end());
                    t.tel = get_data<double>(b.id, "./tel", K_DOUBLE, (size_t)bb.back());
                }
                d.tensors.push_back(std::move(t));
            }
        }
        if (a.lexists(f, "/rOps", H5P_DEFAULT) > 0) {
            Group g(opgroup(f, "/rOps"));
            const int nr = get_attr1<int>(g.id, "nrOps", K_INT);
            for (int i = 0; i < nr; ++i)
                d.nr_rops_groups +=
                    a.lexists(g.id, ("./rOperator_" + std::to_string(i)).c_str(), H5P_DEFAULT) > 0;
        }
    } catch (...) {
        a.fclose(f);
        throw;
    }
    a.fclose(f);
    return d;
}

#else  // built without the HDF5 headers

void write_checkpoint(const std::string&, const Context&, const std::vector<SiteBlocks>&) {
    throw std::runtime_error("checkpoints: this library was built without the HDF5 headers");
}
CheckpointData read_checkpoint(const std::string&) {
    throw std::runtime_error("checkpoints: this library was built without the HDF5 headers");
}

#endif

std::unique_ptr<Context> context_from_checkpoint(const CheckpointData& d) {
    std::vector<int> sgs(d.sym.sgs, d.sym.sgs + d.sym.nrSyms);
    std::vector<int> ts(d.sym.target_state, d.sym.target_state + d.sym.nrSyms);
    auto c = make_context(d.netw, make_model_from_data(d.ham), sgs, ts);
    for (int i = 0; i < c->sym.nrSyms; ++i)
        if (c->sym.sgs[i] != d.sym.sgs[i] || c->sym.target_state[i] != d.sym.target_state[i])
            throw std::runtime_error("checkpoint: symmetries are not in the bookkeeper's order");
    if ((int)d.v_symsecs.size() != d.netw.nr_bonds || (int)d.p_symsecs.size() != d.netw.psites)
        throw std::runtime_error("checkpoint: bookkeeper does not match the network");
    c->v_symsecs = d.v_symsecs;
    c->p_symsecs = d.p_symsecs;
    return c;
}

}  // namespace t3
